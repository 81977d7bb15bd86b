'''What FP32 operands alone cost the density (CPU, numpy; no GPU).

The reduced-precision paths hold the AO values and the MO coefficients in
FP32 (f32) or as three bf16 parts (bf16x6: 24 significant bits, the same
operand precision).  This test rounds the oracle's FP64 AO values and C to
FP32 and forms Phi = C A and rho = sum occ Phi^2 with everything else in
FP64 (exact products of the rounded operands, FP64 sums), on the medium
workload (20 atoms, 400 spherical AOs, 40 MOs) at 4,000 points of its grid.
It pins two facts the precision claims rest on:
  * the norm-wise density error (6e-8 here) is below the 1e-7 target: the
    target is reachable by an FP32-operand path;
  * the largest pointwise relative error (over rho > 1e-10 max rho; 1.3e-7
    here) is above 1e-7: each operand carries up to 2^-24 = 6e-8 relative
    rounding, Phi is bilinear in the operands and rho quadratic in Phi, so
    where the roundings line up the density is off by a few units of 6e-8
    whatever the accumulation precision.
The GPU paths measure rho 3.2e-8 (L2) and 4.9e-7 (pointwise) at full size;
the rest of their pointwise error is the FP32 accumulation over 400 AOs.
'''
import numpy

from orbkit_amd import workloads
from orbkit_amd.tools import get_cart2sph
from oracle import core_oracle as O


def _sample(axes, n, seed):
  X, Y, Z = numpy.meshgrid(*axes, indexing='ij')
  idx = numpy.random.default_rng(seed).choice(X.size, size=n, replace=False)
  return X.ravel()[idx].copy(), Y.ravel()[idx].copy(), Z.ravel()[idx].copy()


def test_fp32_operand_floor():
  qc, axes, _ = workloads.make('medium', npts_side=100)
  x, y, z = _sample(axes, 4000, 21)
  A = O.ao_creator(qc.geo_spec, qc.ao_spec, None, x, y, z, get_cart2sph)   # (400, n), FP64
  C = qc.mo_spec.get_coeffs()
  occ = qc.mo_spec.get_occ()
  phi = C @ A
  rho = occ @ phi ** 2
  r32 = lambda a: a.astype(numpy.float32).astype(numpy.float64)
  phi32 = r32(C) @ r32(A)                    # exact products of rounded operands, FP64 sums
  rho32 = occ @ phi32 ** 2
  mask = rho > 1e-10 * rho.max()
  l2 = numpy.linalg.norm(rho32 - rho) / numpy.linalg.norm(rho)
  pmax = numpy.max(numpy.abs(rho32 - rho)[mask] / rho[mask])
  print('FP32 operands: rho L2 %.2e, pointwise max %.2e' % (l2, pmax))
  assert l2 < 1e-7
  assert pmax > 1e-7
  # and below the bound the operand rounding allows: |d rho| / rho <=
  # 2 max_i |d Phi_i| / |Phi_i|-weighted, with |d Phi| <= 2 u sum |C||A|
  u = 2.0 ** -24
  bound = occ @ (2 * numpy.abs(phi) * (2 * u * (numpy.abs(C) @ numpy.abs(A)) + 1e-300)) + 1e-300
  assert numpy.all(numpy.abs(rho32 - rho)[mask] <= 1.01 * bound[mask])


def test_fp32_rounding_is_what_is_measured():
  '''Control: the same computation with FP64 operands reproduces FP64 to
  rounding level, so the floor above comes from the FP32 operands.'''
  qc, axes, _ = workloads.make('medium', npts_side=100)
  x, y, z = _sample(axes, 500, 22)
  A = O.ao_creator(qc.geo_spec, qc.ao_spec, None, x, y, z, get_cart2sph)
  C = qc.mo_spec.get_coeffs()
  occ = qc.mo_spec.get_occ()
  rho = occ @ (C @ A) ** 2
  rho_b = occ @ numpy.einsum('ik,kp->ip', C, A) ** 2
  assert numpy.linalg.norm(rho_b - rho) / numpy.linalg.norm(rho) < 1e-14
