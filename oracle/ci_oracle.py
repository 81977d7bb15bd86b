'''numpy restatement of orbkit's detCI grid reductions.  TEST INFRASTRUCTURE ONLY.

  rho            orbkit/detci/cy_ci.pyx:70-94 (get_rho): zero-term squares
                 c phi_a^2, then single-excitation products c phi_a phi_b
  jab            orbkit/detci/cy_ci.pyx:143-166 (get_jab):
                 -1/2 c (phi_a d phi_b - phi_b d phi_a) per single excitation
  a_nabla_b      orbkit/detci/cy_ci.pyx:195-229 (get_a_nabla_b; zero terms unused)
  rho_full       orbkit/detci/cy_ci.pyx:96-117 (get_rho_full)
  j_full         orbkit/detci/cy_ci.pyx:119-141 (get_j_full)
  jab_full       orbkit/detci/cy_ci.pyx:168-193 (get_jab_full)

Pinning (tests/test_detci.py): with one zero term per MO (c = occupation,
a = b) rho is the density of the pinned rho_compute oracle; jab is
antisymmetric under exchanging the two orbitals of every single excitation
and equals -1/2 (a_nabla_b - b_nabla_a) with every pair's coefficient; rho_full / j_full with a unit
coefficient for one pair return that pair's density / flux (times the
pair factor); jab_full with ImS of a single pair reduces to jab.
'''
import numpy


def rho(zero, sing, molist):
  out = numpy.zeros(molist.shape[1])
  for iz in range(len(zero[0])):
    for jz in range(len(zero[0][iz])):
      sta = zero[1][iz][jz]
      out += zero[0][iz][jz] * molist[sta] * molist[sta]
  for ks in range(len(sing[0])):
    sta, stb = sing[1][ks][0], sing[1][ks][1]
    out += sing[0][ks] * molist[sta] * molist[stb]
  return out


def jab(zero, sing, molist, molistdrv):
  out = numpy.zeros((3, molist.shape[1]))
  for ks in range(len(sing[0])):
    sta, stb = sing[1][ks][0], sing[1][ks][1]
    c = sing[0][ks]
    for d in range(3):
      out[d] -= 0.5 * (c * (molist[sta] * molistdrv[d, stb] - molist[stb] * molistdrv[d, sta]))
  return out


def a_nabla_b(zero, sing, molist, molistdrv):
  '''single-excitation terms only (the zero terms are not used)'''
  out = numpy.zeros((3, molist.shape[1]))
  for ks in range(len(sing[0])):
    sta, stb = sing[1][ks][0], sing[1][ks][1]
    for d in range(3):
      out[d] += sing[0][ks] * molist[sta] * molistdrv[d, stb]
  return out


def rho_full(ReS, rho):
  nt, nstate = ReS.shape[0], ReS.shape[1]
  out = numpy.zeros((nt, rho.shape[1]))
  for t in range(nt):
    count = 0
    for n in range(nstate):
      for m in range(n, nstate):
        out[t] += (ReS[t, m, m] if m == n else 2.0 * ReS[t, m, n]) * rho[count]
        count += 1
  return out


def j_full(ImS, j):
  nt, nstate = ImS.shape[0], ImS.shape[1]
  out = numpy.zeros((nt, 3, j.shape[2]))
  for t in range(nt):
    count = 0
    for n in range(nstate):
      for m in range(n, nstate):
        if m != n:
          out[t] -= 2.0 * ImS[t, n, m] * j[count]
        count += 1
  return out


def jab_full(ImS, chi, nabla_chi, mu):
  f = 1.0 / mu
  out = numpy.zeros((nabla_chi.shape[0], chi.shape[1]))
  for c in range(nabla_chi.shape[0]):
    for n in range(ImS.shape[0]):
      for m in range(n):
        out[c] += f * ImS[n, m] * (chi[n] * nabla_chi[c, m] - chi[m] * nabla_chi[c, n])
  return out
