'''numpy restatement of orbkit's grid analysis helpers.  TEST INFRASTRUCTURE ONLY.

  gross_atomic_density   orbkit/extras.py:306-385: per selected atom the MOs
                         contracted over that atom's AO rows only, times the
                         full MOs, occupation-weighted (loop order as orbkit)
  mo_matrix              orbkit/core.py:841-941: bra MOs (no derivative)
                         times ket MOs (each requested derivative), all pairs
  jmo                    orbkit/extras.py:441-492: -1/2 (m[:, i, j] - m[:, j, i])
  cube_text              orbkit/output/cube.py:21-96: header and value text

AO rows are attributed to atoms through the output rows' contractions (for
spherical bases orbkit's own loop counts Cartesian rows per contraction,
orbkit/extras.py:361-368, which misaligns them; the restatement follows the
rows actually produced).

Pinning (tests/test_extras.py): the gross atomic densities of all atoms add
up to the density of the pinned rho_compute oracle; for a one-atom molecule
the gross density is the density; sum_n occ_n mo_matrix[0, n, n] is the
density; jmo is antisymmetric in (i, j) and vanishes for i = j; the cube
text reproduces Gaussian cubegen value blocks (tests/golden) byte for byte.
'''
import numpy

from oracle import grad_oracle


def row_atoms(qc):
  '''Atom (0-based) of each output AO row.'''
  return grad_oracle.row_map(qc)[1]


def gross_atomic_density(qc, index, ao_list, mo_list):
  '''orbkit/extras.py:355-377 for atoms `index` (0-based geo indices).'''
  C, occ = qc.mo_spec.get_coeffs(), qc.mo_spec.get_occ()
  ra = row_atoms(qc)
  N = mo_list.shape[1:]
  out = []
  for a in index:
    rows = numpy.nonzero(ra == a)[0]
    rho = numpy.zeros(N)
    for i in range(C.shape[0]):
      mo_info = numpy.zeros(N)
      for jj in rows:
        mo_info += C[i, jj] * ao_list[jj]
      rho += occ[i] * mo_list[i] * mo_info
    out.append(rho)
  return out


def mo_matrix(mo_bra, mo_ket):
  '''orbkit/core.py:935-939: mo_matrix[:, n, m] = mo_bra[:, n] * mo_ket[:, m]
  with mo_bra (1, NMO_a) + N and mo_ket (NDRV, NMO_b) + N.'''
  shape = (mo_ket.shape[0], mo_bra.shape[1]) + mo_ket.shape[1:]
  out = numpy.zeros(shape)
  for n in range(mo_bra.shape[1]):
    for m in range(mo_ket.shape[1]):
      out[:, n, m] = mo_bra[:, n] * mo_ket[:, m]
  return out


def jmo(mo_mat, indices):
  '''orbkit/extras.py:482-485.'''
  out = numpy.zeros((mo_mat.shape[0], len(indices)) + mo_mat.shape[3:])
  for n, (i, j) in enumerate(indices):
    out[:, n] = -0.5 * (mo_mat[:, i, j] - mo_mat[:, j, i])
  return out


def cube_text(data, geo_info, geo_spec, min_, N_, delta_, comments='', labels=None):
  '''orbkit/output/cube.py:21-96 as one string (string-building loops as
  orbkit, grid passed explicitly).  Pinned by the Gaussian cubegen files in
  tests/golden (their value text reproduced byte for byte) and by reading the
  values back.'''
  data = numpy.array(data)
  if data.ndim == 3:
    data = data[numpy.newaxis]
  string = 'orbkit calculation\n'
  string += ' %(f)s\n' % {'f': comments}
  string += ('%(at)d' % {'at': (-1) ** (labels is not None) * len(geo_info)}).rjust(5)
  for ii in range(3):
    string += ('%(min)0.6f' % {'min': min_[ii]}).rjust(12)
  if len(data) > 1:
    string += ('%d' % len(data)).rjust(12)
  for ii in range(3):
    string += '\n'
    string += ('%(N)d' % {'N': N_[ii]}).rjust(5)
    for jj in range(3):
      if jj == ii:
        string += ('%(dr)0.6f' % {'dr': delta_[ii]}).rjust(12)
      else:
        string += ('%(dr)0.6f' % {'dr': 0}).rjust(12)
  string += '\n'
  for ii in range(len(geo_info)):
    string += ('%(N)d' % {'N': round(float(geo_info[ii][2]))}).rjust(5)
    string += ('%(ch)0.6f' % {'ch': float(geo_info[ii][1])}).rjust(12)
    for jj in range(3):
      string += ('%(r)0.6f' % {'r': geo_spec[ii][jj]}).rjust(12)
    string += '\n'
  if labels is not None:
    string += ('%(N)d' % {'N': len(data)}).rjust(5)
    c = 0
    for i, j in enumerate(labels):
      c += 1
      string += str(j).rjust(5)
      if c % 9 == 8:
        string += '\n'
    string += '\n'
  for rr in range(N_[0]):
    for ss in range(N_[1]):
      c = 0
      for tt in range(N_[2]):
        for dd in data[:, rr, ss, tt]:
          string += ('%(data).5E' % {'data': dd}).rjust(13)
          if c % 6 == 5:
            string += '\n'
          c += 1
      string += '\n'
  return string
