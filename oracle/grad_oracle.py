'''FP64 numpy gradients of the electron density.  TEST INFRASTRUCTURE ONLY.

orbkit has no derivative code.  The forward here is orbkit's density path
restated by core_oracle (AO rows: orbkit/core.py:38-105 via the C
restatement of c_lcreator; 'N' row factor: orbkit/core.py:98-100;
cartesian2spherical: orbkit/core.py:135-176; MOs and occupation-weighted
squares: orbkit/core.py:179-308), with exponents and nuclear positions as
explicit parameters.  vjp() gives the gradient of a scalar loss through
rho by the chain rule on AO values:

  Phi = C A,  B = 2 occ g Phi,  dL/dC = B A^T,  G = C^T B,
  dL/dR_a   = -sum_r sum_{rows mu on atom a} G_mu grad A_mu
  dL/dalpha_k = sum_r sum_mu G_mu dA_mu/dalpha_k, with dA/dalpha_k the
              single-primitive AO term of primitive k times
              ((2l+3)/(4 alpha_k) - |r - R|^2)  (the first part only for
              primitives orbkit normalises, orbkit/c_support.c:177-188)

Pinning (tests/test_oracle.py): central finite differences of loss(forward)
for every exponent and position component and sampled C entries, the
homogeneity sum C dL/dC = 2 sum g rho, and translation invariance
sum_a dL/dR_a = -sum_r g grad rho (grad rho from the forward's own
derivative AOs).
'''
import numpy

from oracle import core_oracle


def _flat(qc, alpha=None, geo=None):
  f = core_oracle.flat_basis(qc.geo_spec, qc.ao_spec)
  f['prims'] = numpy.array(f['prims'], dtype=numpy.float64, copy=True)
  if alpha is not None:
    f['prims'][:, 0] = numpy.asarray(alpha, dtype=numpy.float64)
  if geo is not None:
    f['geo'] = numpy.ascontiguousarray(geo, dtype=numpy.float64).reshape(-1, 3)
  return f


def row_map(qc):
  '''(Tm, row_atom): output rows = Tm @ (N * Cartesian rows); atom of each output row.'''
  from orbkit_amd.tools import get_cart2sph
  ao_spec = qc.ao_spec
  ncart = int(numpy.sum(ao_spec.get_nlxlylz_per_cont()))
  Nf = numpy.ones(ncart)
  if 'N' in ao_spec[0]:
    Nf = numpy.asarray(ao_spec[0]['N'], dtype=numpy.float64).reshape(-1)
  eye = numpy.diag(Nf)
  cont_atom = ao_spec.get_assign_cont_to_atoms()
  if ao_spec.spherical:
    Tm = core_oracle.cartesian2spherical(eye, ao_spec, get_cart2sph)
    row_atom = numpy.array([cont_atom[j0] for j0, _ in ao_spec.get_old_ao_spherical()])
  else:
    Tm = eye
    row_atom = numpy.asarray(cont_atom)[ao_spec.get_assign_lxlylz_to_cont()]
  return Tm, row_atom


def aos(qc, x, y, z, code=0, alpha=None, geo=None, Tm=None):
  f = _flat(qc, alpha, geo)
  if Tm is None:
    Tm, _ = row_map(qc)
  return Tm @ core_oracle.aocreator(f, x, y, z, code)


def density(qc, x, y, z, C=None, occ=None, alpha=None, geo=None):
  C = qc.mo_spec.get_coeffs() if C is None else numpy.asarray(C)
  occ = qc.mo_spec.get_occ() if occ is None else numpy.asarray(occ)
  Phi = C @ aos(qc, x, y, z, 0, alpha, geo)
  return occ @ Phi ** 2


def vjp(qc, x, y, z, g, C=None, occ=None, alpha=None, geo=None):
  '''(rho, dL/dC, dL/dalpha, dL/dgeo) for dL/drho = g.'''
  C = qc.mo_spec.get_coeffs() if C is None else numpy.asarray(C)
  occ = qc.mo_spec.get_occ() if occ is None else numpy.asarray(occ)
  f = _flat(qc, alpha, geo)
  Tm, row_atom = row_map(qc)
  A = Tm @ core_oracle.aocreator(f, x, y, z, 0)
  dA = [Tm @ core_oracle.aocreator(f, x, y, z, c) for c in (1, 2, 3)]
  Phi = C @ A
  rho = occ @ Phi ** 2
  B = 2.0 * occ[:, None] * g[None, :] * Phi
  dC = B @ A.T
  G = C.T @ B
  natoms = len(f['geo'])
  dgeo = numpy.zeros((natoms, 3))
  for d in range(3):
    per_row = numpy.sum(G * dA[d], axis=1)
    for a in range(natoms):
      dgeo[a, d] = -numpy.sum(per_row[row_atom == a])
  # exponents: single-primitive AO terms of each contraction
  ncart = f['ncart']
  nprim = f['nprim']
  lxlylz = numpy.asarray(f['lxlylz'])
  dalpha = numpy.zeros(len(f['prims']))
  c0 = p0 = 0
  for j in range(len(ncart)):
    nc, npj = int(ncart[j]), int(nprim[j])
    at = int(f['atoms'][j])
    l = int(numpy.sum(lxlylz[c0]))
    R = f['geo'][at]
    rr = (x - R[0]) ** 2 + (y - R[1]) ** 2 + (z - R[2]) ** 2
    nfac = 0.0 if f['normalized'] else (2.0 * l + 3.0) / 4.0
    Gj = Tm[:, c0:c0 + nc].T @ G          # dL/d(Cartesian rows of contraction j)
    for p in range(npj):
      k = p0 + p
      fj = dict(lxlylz=lxlylz[c0:c0 + nc], ncart=numpy.array([nc]), prims=f['prims'][k:k + 1],
                nprim=numpy.array([1]), atoms=numpy.array([0]), normalized=f['normalized'],
                geo=R.reshape(1, 3))
      term = core_oracle.aocreator(fj, x, y, z, 0)      # (nc, n)
      dalpha[k] = numpy.sum(Gj * term * (nfac / f['prims'][k, 0] - rr)[None, :])
    c0 += nc
    p0 += npj
  return rho, dC, dalpha, dgeo
