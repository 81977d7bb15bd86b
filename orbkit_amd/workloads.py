'''Seeded synthetic molecules for benchmarks and tests.

Each workload is a qc object (geo_spec, ao_spec, mo_spec) plus a regular grid:
atoms uniform in a cube (pairs closer than 1 Angstrom rejected), every atom
carrying a fixed shell pattern with 3 primitives per shell, exponents
log-uniform in a range, contraction coefficients uniform in (0, 1) scaled so
the contracted function has unit self-overlap under orbkit's primitive
normalisation, MO coefficients standard normal then Loewdin-orthonormalised
in the AO overlap metric, every MO doubly occupied, and a regular grid over
the atoms' bounding box padded by 5 bohr.

  small     3 atoms, 4 A box, Cartesian s/p/d, 25 AOs, 5 MOs, 50^3 points
  medium    20 atoms, 12 A box, spherical [s,s,p,p,d,f], 400 AOs, 40 MOs, 100^3
  gradient  the medium basis on 80^3 points (loss and gradient workload)
  batch     256 molecules, 5-60 atoms of [s,s,p,d] (spherical), 8-30 MOs,
            grid sides 32-80
  large     150 atoms [s,s,p,p,d] + 50 atoms [s,s,p,p,d,f], 30 A box,
            2950 AOs, 400 MOs, 128^3
'''
from math import comb

import numpy

from .orbitals import AOClass, MOClass
from .qcinfo import QCinfo
from .tools import exp as cart_exp, lquant
from .units import aa_to_a0
from .read.overlap import dfact, self_overlap
from .basis import spherical_matrix


def _atoms(rng, n, box_aa):
  pts = []
  while len(pts) < n:
    p = rng.uniform(0, box_aa, 3)
    if all(numpy.linalg.norm(p - q) >= 1.0 for q in pts):
      pts.append(p)
  return numpy.array(pts) * aa_to_a0


def _basis(rng, geo, patterns, lo, hi, spherical, nprim=3):
  ao = AOClass()
  for a, pat in enumerate(patterns):
    for t in pat:
      l = lquant[t]
      al = numpy.sort(numpy.exp(rng.uniform(numpy.log(lo), numpy.log(hi), nprim)))[::-1]
      c = rng.uniform(0, 1, nprim)
      co = numpy.stack([al, c], axis=1)
      co[:, 1] /= numpy.sqrt(self_overlap((l, 0, 0), co, False))
      ao.append({'atom': a, 'type': t, 'pnum': nprim, 'coeffs': co})
  if spherical:
    ao.set_lm_dict(p=[1, 0])
  ao.update()
  return ao


def _int1d(n, p):
  '''int t^n exp(-p t^2) dt over R (vectorised over p).'''
  if n % 2:
    return numpy.zeros_like(p)
  return dfact(n - 1) / (2 * p) ** (n // 2) * numpy.sqrt(numpy.pi / p)


def cartesian_overlap(geo, ao_spec):
  '''Analytic overlap of the Cartesian AO rows under orbkit's primitive
  normalisation (Gaussian product theorem + binomial expansion).'''
  lxlylz = ao_spec.get_lxlylz()
  ncart = ao_spec.get_nlxlylz_per_cont()
  nprim = ao_spec.get_nprim_per_cont()
  pc = ao_spec.get_prim_coeffs()
  atoms = ao_spec.get_assign_cont_to_atoms()
  nc = len(ncart)
  row0 = numpy.concatenate([[0], numpy.cumsum(ncart)])
  p0 = numpy.concatenate([[0], numpy.cumsum(nprim)])
  ls = numpy.array([lxlylz[row0[c]].sum() for c in range(nc)])
  P = int(nprim.max())
  al = numpy.zeros((nc, P))
  cn = numpy.zeros((nc, P))
  for c in range(nc):
    k = nprim[c]
    al[c, :k] = pc[p0[c]:p0[c] + k, 0]
    l = ls[c]
    cn[c, :k] = pc[p0[c]:p0[c] + k, 1] * (2 / numpy.pi) ** 0.75 * 2 ** l * al[c, :k] ** ((2 * l + 3) / 4)
    al[c, k:] = 1.0
  S = numpy.zeros((int(ncart.sum()),) * 2)
  cen = geo[atoms]
  for la in range(ls.max() + 1):
    for lb in range(ls.max() + 1):
      ia = numpy.nonzero(ls == la)[0]
      ib = numpy.nonzero(ls == lb)[0]
      if not len(ia) or not len(ib):
        continue
      A = cen[ia][:, None, None, None, :]
      B = cen[ib][None, :, None, None, :]
      a = al[ia][:, None, :, None]
      b = al[ib][None, :, None, :]
      p = a + b
      mu = a * b / p
      AB2 = ((cen[ia][:, None, :] - cen[ib][None, :, :]) ** 2).sum(-1)[:, :, None, None]
      E = numpy.exp(-mu * AB2)
      Pc = (a[..., None] * A + b[..., None] * B) / p[..., None]
      PA, PB = Pc - A, Pc - B
      # 1-D overlaps I[axis][i][j] for powers i <= la, j <= lb
      I = numpy.zeros((3, la + 1, lb + 1) + p.shape)
      for ax in range(3):
        for i in range(la + 1):
          for j in range(lb + 1):
            s = 0.0
            for u in range(i + 1):
              for v in range(j + 1):
                if (u + v) % 2:
                  continue
                s = s + comb(i, u) * comb(j, v) * PA[..., ax] ** (i - u) * PB[..., ax] ** (j - v) * _int1d(u + v, p)
            I[ax, i, j] = s
      wa = cn[ia][:, None, :, None]
      wb = cn[ib][None, :, None, :]
      W = wa * wb * E
      for ca, ea in enumerate(cart_exp[la]):
        da = numpy.sqrt(dfact(2 * ea[0] - 1) * dfact(2 * ea[1] - 1) * dfact(2 * ea[2] - 1))
        for cb, eb in enumerate(cart_exp[lb]):
          db = numpy.sqrt(dfact(2 * eb[0] - 1) * dfact(2 * eb[1] - 1) * dfact(2 * eb[2] - 1))
          val = (W * I[0, ea[0], eb[0]] * I[1, ea[1], eb[1]] * I[2, ea[2], eb[2]]).sum(axis=(2, 3)) / (da * db)
          S[numpy.ix_(row0[ia] + ca, row0[ib] + cb)] = val
  return S


def ao_overlap(geo, ao_spec):
  S = cartesian_overlap(geo, ao_spec)
  if ao_spec.spherical:
    T = spherical_matrix(ao_spec)
    S = T @ S @ T.T
  return S


def _mos(rng, S, nmo):
  nao = S.shape[0]
  C0 = rng.standard_normal((nmo, nao))
  M = C0 @ S @ C0.T
  w, V = numpy.linalg.eigh(M)
  C = (V @ numpy.diag(w ** -0.5) @ V.T) @ C0
  mo = MOClass([{'coeffs': C[i], 'occ_num': 2.0, 'energy': -1.0 + 0.01 * i, 'sym': '%d.1' % (i + 1)}
                for i in range(nmo)])
  return mo


def _grid(geo, n, pad=5.0):
  lo = geo.min(axis=0) - pad
  hi = geo.max(axis=0) + pad
  return [numpy.linspace(lo[i], hi[i], n) for i in range(3)]


def make(name, seed=None, npts_side=None, nmo=None):
  '''Returns (qc, (x, y, z) regular axes, info dict).  ``nmo`` overrides the
  workload's number of (orthonormal, doubly occupied) MOs.'''
  nmo_override = nmo
  seeds = {'small': 1, 'medium': 2, 'gradient': 3, 'large': 5}
  rng = numpy.random.default_rng(seeds[name] if seed is None else seed)
  if name == 'small':
    geo = _atoms(rng, 3, 4.0)
    ao = _basis(rng, geo, [['s', 'p', 'd'], ['s', 'p', 'd'], ['s', 's', 'p']], 0.1, 5000.0, False)
    nmo, side, slice_length = 5, 50, 1e4
  elif name in ('medium', 'gradient'):
    if name == 'gradient':
      rng = numpy.random.default_rng(2 if seed is None else seed)
    geo = _atoms(rng, 20, 12.0)
    ao = _basis(rng, geo, [['s', 's', 'p', 'p', 'd', 'f']] * 20, 0.05, 1e5, True)
    nmo, side, slice_length = 40, (100 if name == 'medium' else 80), 1e4
  elif name == 'large':
    geo = _atoms(rng, 200, 30.0)
    pats = [['s', 's', 'p', 'p', 'd']] * 150 + [['s', 's', 'p', 'p', 'd', 'f']] * 50
    ao = _basis(rng, geo, pats, 0.05, 1e5, True)
    nmo, side, slice_length = 400, 128, 65536
  else:
    raise ValueError('unknown workload %r' % name)
  if npts_side is not None:
    side = npts_side
  if nmo_override is not None:
    nmo = int(nmo_override)
  S = ao_overlap(geo, ao)
  mo = _mos(rng, S, nmo)
  qc = QCinfo()
  qc.geo_spec = geo
  qc.geo_info = numpy.array([['X', str(i + 1), '0'] for i in range(len(geo))])
  qc.ao_spec = ao
  qc.mo_spec = mo
  axes = _grid(geo, side)
  info = {'workload': name, 'nao': ao.get_ao_num(), 'nmo': nmo, 'natoms': len(geo),
          'nprim': int(ao.get_nprim_per_cont().sum()), 'side': side, 'slice_length': slice_length}
  return qc, axes, info


def make_batch(n_mol=256, seed=4, side_range=(32, 80)):
  '''List of (qc, axes) for the batched workload.'''
  rng = numpy.random.default_rng(seed)
  out = []
  for m in range(n_mol):
    nat = int(rng.integers(5, 61))
    geo = _atoms(rng, nat, 4.0 + 0.25 * nat)
    ao = _basis(rng, geo, [['s', 's', 'p', 'd']] * nat, 0.05, 1e5, True)
    nmo = int(rng.integers(8, 31))
    S = ao_overlap(geo, ao)
    mo = _mos(rng, S, nmo)
    qc = QCinfo()
    qc.geo_spec, qc.ao_spec, qc.mo_spec = geo, ao, mo
    side = int(rng.integers(side_range[0], side_range[1] + 1))
    out.append((qc, _grid(geo, side)))
  return out
