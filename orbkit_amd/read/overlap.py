'''Closed-form same-centre overlaps used by the readers' normalisation check.

For a Cartesian Gaussian component x^lx y^ly z^lz with primitives (a_p, c_p)
on one centre, <g_p|g_q> = prod_i (2l_i-1)!!/(2s)^{l_i} * (pi/s)^{3/2}
with s = a_p + a_q.  Under orbkit's primitive normalisation
N_p = (2/pi)^{3/4} 2^l a_p^{(2l+3)/4} / sqrt(prod_i (2l_i-1)!!)
(orbkit/c_support.c:177-188) this becomes (2 sqrt(a_p a_q)/s)^{l+3/2}, the
same for every component of the shell.  These are the diagonal elements of
orbkit's analytic overlap matrix (orbkit/analytical_integrals.py:35-118),
which orbkit/read/molden.py:377-398 uses to decide on renormalisation.
'''
import numpy


def dfact(n):
  r = 1
  while n > 0:
    r *= n
    n -= 2
  return r


def self_overlap(lxlylz, coeffs, normalized):
  '''Diagonal overlap of one Cartesian component ``lxlylz`` of a contraction
  with primitive table ``coeffs`` (pnum x 2: exponent, coefficient).'''
  lx, ly, lz = (int(v) for v in lxlylz)
  l = lx + ly + lz
  a = numpy.asarray(coeffs, dtype=float)[:, 0]
  c = numpy.asarray(coeffs, dtype=float)[:, 1]
  s = a[:, None] + a[None, :]
  if normalized:
    df = dfact(2 * lx - 1) * dfact(2 * ly - 1) * dfact(2 * lz - 1)
    prim = df / (2 * s) ** l * (numpy.pi / s) ** 1.5
  else:
    prim = (2 * numpy.sqrt(a[:, None] * a[None, :]) / s) ** (l + 1.5)
  return float(c @ prim @ c)


def self_overlap_rows(ao_spec):
  '''Self-overlap of every Cartesian AO row of ``ao_spec``.'''
  lxlylz = ao_spec.get_lxlylz()
  per = ao_spec.get_nlxlylz_per_cont()
  normalized = ao_spec.get_normalized()
  out, r = [], 0
  for ic, n in enumerate(per):
    co = ao_spec[ic]['coeffs']
    for k in range(n):
      out.append(self_overlap(lxlylz[r + k], co, normalized))
    r += n
  return numpy.array(out)


def omitted_cca_norm(lxlylz):
  '''sqrt(prod_i (2l_i-1)!! / (2l-1)!!) per Cartesian row: the factor the
  CCA convention leaves out of Cartesian functions.'''
  out = []
  for lx, ly, lz in numpy.asarray(lxlylz):
    out.append(numpy.sqrt(dfact(2 * lx - 1) * dfact(2 * ly - 1) * dfact(2 * lz - 1)
                          / dfact(2 * (lx + ly + lz) - 1)))
  return numpy.array(out)
