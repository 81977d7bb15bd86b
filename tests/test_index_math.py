'''The fused and batch kernels' index arithmetic, replayed on the CPU with
bounds checks (no GPU).  Restates, line for line in meaning, the device code
in orbkit_amd/csrc/fused_tile.h and rho_fused.hip:

  * fetch_c / put_c / read_c: the split C operand of one chunk, loaded by the
    256 threads of a workgroup from the coefficient planes
    [plane][mpad][kpad] and staged in LDS as 80-byte rows (CStage);
  * find_item: the batch kernel's binary search from a tile index to its
    molecule over the tile prefix sums;
  * brick_of / tile_point: a tile's 64 points in a 2^tlx x 2^tly x 2^tlz
    brick of a regular grid, clamped at the grid's faces.

Every global read must stay inside the planes, every LDS write inside the
stage, every contraction read on a written slot, every tile must map to the
molecule whose range holds it, and the active lanes of all bricks must cover
each grid point exactly once, for the shapes the GPU tests and the
batched workload use (one to four MO tiles, two or three bf16 planes, every
brick shape the tuner tries, grid sides 12..80).
'''
import numpy
import pytest

KC = 32


def cstage(nplanes, mt):
  data = nplanes * mt * 16 * 4
  return data, nplanes * mt * 16 * 5, (data + 255) // 256


@pytest.mark.parametrize('nplanes', [2, 3])
@pytest.mark.parametrize('mt', [1, 2, 3, 4])
def test_c_stage_bounds(nplanes, mt):
  data, slots, rounds = cstage(nplanes, mt)
  mpad = mt * 16
  for kpad in (32, 64, 416):
    nplane_elems = nplanes * mpad * kpad
    written = numpy.zeros(slots, dtype=bool)
    for k0 in range(0, kpad, KC):
      written[:] = False
      for t in range(rounds):
        for thread in range(256):
          # fetch_c: unconditional load, clamped to the last data slot
          d = min(t * 256 + thread, data - 1)
          r, c = d >> 2, d & 3
          q, m = r // (mt * 16), r - (r // (mt * 16)) * (mt * 16)
          assert 0 <= q < nplanes and 0 <= m < mpad
          first = (q * mpad + m) * kpad + k0 + 8 * c
          assert 0 <= first and first + 7 < nplane_elems
          assert k0 + 8 * c + 7 < kpad             # inside this MO row
          # put_c: only data slots are written
          d = t * 256 + thread
          if d < data:
            slot = (d >> 2) * 5 + (d & 3)
            assert slot < slots and not written[slot]
            written[slot] = True
      # read_c: every fragment read is a written slot
      for q in range(nplanes):
        for m_t in range(mt):
          for n16 in range(16):
            for g in range(4):
              slot = ((q * mt + m_t) * 16 + n16) * 5 + g
              assert slot < slots and written[slot]
      assert written.sum() == data


def find_item(start, n, t):
  lo, hi = 0, n - 1
  while lo < hi:
    mid = (lo + hi + 1) >> 1
    if start[mid] <= t:
      lo = mid
    else:
      hi = mid - 1
  return lo


def test_find_item_maps_every_tile():
  rng = numpy.random.default_rng(3)
  for n in (1, 2, 10, 256):
    tiles = rng.integers(1, 1000, size=n)
    start = numpy.concatenate([[0], numpy.cumsum(tiles)])
    for t in range(0, int(start[-1]), max(1, int(start[-1]) // 5000)):
      m = find_item(start, n, t)
      assert start[m] <= t < start[m + 1]
    for m in range(n):   # both ends of every molecule's range
      assert find_item(start, n, int(start[m])) == m
      assert find_item(start, n, int(start[m + 1]) - 1) == m


@pytest.mark.parametrize('shape', [(2, 2, 2), (1, 2, 3), (0, 3, 3), (0, 0, 6)])
@pytest.mark.parametrize('sides', [(12, 13, 14), (19, 19, 19), (32, 47, 80)])
def test_bricks_cover_the_grid_once(shape, sides):
  tlx, tly, tlz = shape
  nx, ny, nz = sides
  nbx, nby, nbz = -(-nx >> tlx), -(-ny >> tly), -(-nz >> tlz)
  ntile = nbx * nby * nbz
  assert ntile < 2 ** 31
  count = numpy.zeros(nx * ny * nz, dtype=int)
  lane = numpy.arange(64)
  for tile in range(ntile):
    t = tile                          # brick_of (32-bit arithmetic)
    bz = t % nbz
    t //= nbz
    by = t % nby
    bx = t // nby
    ix = (bx << tlx) + (lane >> (tly + tlz))
    iy = (by << tly) + ((lane >> tlz) & ((1 << tly) - 1))
    iz = (bz << tlz) + (lane & ((1 << tlz) - 1))
    act = (ix < nx) & (iy < ny) & (iz < nz)
    ix, iy, iz = numpy.minimum(ix, nx - 1), numpy.minimum(iy, ny - 1), numpy.minimum(iz, nz - 1)
    p = (ix * ny + iy) * nz + iz
    assert p.min() >= 0 and p.max() < nx * ny * nz     # clamped lanes read inside the grid
    numpy.add.at(count, p[act], 1)
  assert numpy.all(count == 1)


def _group_sum(alpha, d, rcut2, d2, rr):
  '''The fused kernels' primitive loop (fused_tile.h eval_shell_chunk):
  groups of four; the number of primitives of a group that reach the tile
  (d2 < rcut2) is counted and that many LEADING primitives are summed; a
  group that is not fully live ends the shell.'''
  s = 0.0
  for q in range(0, len(alpha), 4):
    nl = int(numpy.sum(d2 < rcut2[q:q + 4]))
    for j in range(nl):
      s += d[q + j] * numpy.exp(-alpha[q + j] * rr)
    if nl < 4:
      break
  return s


def _fused_groups(alpha, d, rcut2):
  '''prepare_fused_tables: a shell's primitives by screening radius, largest
  first (stable), in groups of four padded with never-live slots.'''
  order = sorted(range(len(alpha)), key=lambda k: -rcut2[k])
  pad = (-len(alpha)) % 4
  return (numpy.concatenate([alpha[order], numpy.zeros(pad)]), numpy.concatenate([d[order], numpy.zeros(pad)]),
          numpy.concatenate([rcut2[order], -numpy.ones(pad)]))


def test_primitive_groups_need_radius_order():
  """With the primitives sorted by screening radius the group loop sums
  exactly the primitives that reach the tile (the per-primitive screened
  sum, which is what the FP64 AO kernel computes in file order).  In file
  order, exponents listed tightest first (as the workloads and most basis
  files list them), the same loop drops live primitives whenever a tight one
  is screened out: the mechanism of the error a four-at-a-time FP64 AO
  kernel variant showed on the medium basis (taken out, never committed),
  whose primitive table is in file order.  The kept kernels build the
  sorted groups on the host."""
  rng = numpy.random.default_rng(8)
  dropped = 0
  for trial in range(2000):
    n = int(rng.integers(1, 9))
    alpha = numpy.sort(numpy.exp(rng.uniform(numpy.log(0.05), numpy.log(1e5), n)))[::-1]   # tightest first
    d = rng.uniform(0.1, 1.0, n)
    rcut2 = 92.1 / alpha                      # exp(-alpha r^2) below 1e-40 beyond r^2 = 92.1 / alpha
    d2 = float(rng.uniform(0, 2 * rcut2.max()))
    rr = d2 + float(rng.uniform(0, 1))
    exact = sum(d[k] * numpy.exp(-alpha[k] * rr) for k in range(n) if d2 < rcut2[k])
    assert _group_sum(*_fused_groups(alpha, d, rcut2), d2, rr) == pytest.approx(exact, rel=1e-15, abs=0)
    if abs(_group_sum(alpha, d, rcut2, d2, rr) - exact) > 1e-12 * abs(exact):
      dropped += 1
  assert dropped > 100   # file order loses live primitives in many configurations


def _mo_group(nsets, nmt):
  '''MO tiles per group of the FP64 density contraction (rho_gemm_impl.h
  mo_group): at most 24 accumulator tiles, at most four MO tiles.'''
  mg = max(1, min(4, 24 // max(nsets, 1)))
  return min(nmt, mg)


def _rho_f64_passes(nsc, nmo):
  '''The passes each wave of one k_rho_f64 workgroup runs (launch_rho,
  launch_rho_full and the kernel's share arithmetic): a list per wave of
  (first tile, tiles run).'''
  nmt = (nmo + 15) // 16
  mg = _mo_group(nsc, nmt)
  ng = (nmt + mg - 1) // mg
  wg = ng >= 2
  exact = mg == 4 and mg * nsc <= 16
  w8 = 8 if exact else ng
  nwave = (ng if ng <= 4 else w8 if ng <= 8 else 4) if wg else 1
  assert 1 <= nwave <= 8                       # __launch_bounds__(512), red[8]
  waves = []
  for wave in range(nwave):
    passes = []
    if exact:
      t_beg = nmt * wave // nwave if wg else 0
      t_end = nmt * (wave + 1) // nwave if wg else nmt
      t0 = t_beg
      while t0 < t_end:
        nt = min(mg, t_end - t0)
        passes.append((t0, nt))
        t0 += mg
    else:
      g = wave if wg else 0
      while g < ng:
        passes.append((g * mg, mg))
        g += nwave if wg else 1
    waves.append(passes)
  return nmt, mg, waves


@pytest.mark.parametrize('nsc', [1, 4, 7, 10])
def test_rho_f64_tile_shares_cover_each_mo_once(nsc):
  '''Every 16-MO tile is contracted by exactly one pass of one wave, every
  pass runs at most MG tiles, and the tiles a pass runs past the last MO are
  only those of the last group (masked by m < nmo in the kernel).'''
  for nmo in range(1, 1200):
    nmt, mg, waves = _rho_f64_passes(nsc, nmo)
    seen = numpy.zeros(nmt + 4, dtype=int)
    for passes in waves:
      for t0, nt in passes:
        assert 1 <= nt <= mg
        seen[t0:t0 + nt] += 1
    assert (seen[:nmt] == 1).all(), (nsc, nmo)
    assert seen[nmt:].sum() <= mg - 1, (nsc, nmo)
    assert all(t0 < nmt for passes in waves for t0, nt in passes)
    if mg == 4 and mg * nsc <= 16:
      # exact passes: none past the last tile, shares differ by one tile
      assert all(t0 + nt <= nmt for passes in waves for t0, nt in passes)
      shares = [sum(nt for _, nt in passes) for passes in waves]
      assert max(shares) - min(shares) <= 1, (nsc, nmo, shares)
