'''Compiles liborbkit_amd.so for gfx950 in-tree (hipcc, no JIT cache).

    python -m orbkit_amd.build
'''
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OBJ = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'liborbkit_amd.so')
INCLUDE = os.path.join(os.path.dirname(HERE), 'include')

SOURCES = ['abi.cpp', 'ao_eval.hip', 'rho_gemm.hip', 'rho_gemm_m0.hip', 'rho_gemm_m1.hip', 'rho_fused.hip', 'rho_fused_f64.hip', 'rho_backward.hip', 'rho_big.hip', 'electrons.hip', 'extras.hip', 'cube.cpp', 'detci.hip']

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=gfx950', '-O3', '-fPIC', '-std=c++17', '-x', 'hip',
         # packed FP32 ops stay off: the compiler would otherwise form
         # v_pk_{add,mul,fma}_f32 from plain float code
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-munsafe-fp-atomics', '-I', INCLUDE, '-I', CSRC]

# per-source additions:
#   rho_fused.hip, rho_fused_f64.hip: their only atomic is a work counter
#     taken by one lane, which the atomic optimizer would wrap in a
#     wave-reduction sequence per claim;
#   rho_gemm_m1.hip, rho_fused_f64.hip: FP64 MFMAs in the VGPR form (by
#     default each K step copied the accumulators into AGPRs and back around
#     every MFMA).
EXTRA = {'rho_fused.hip': ['-mllvm', '-amdgpu-atomic-optimizer-strategy=None'],
         'rho_fused_f64.hip': ['-mllvm', '-amdgpu-atomic-optimizer-strategy=None', '-mllvm', '-amdgpu-mfma-vgpr-form'],
         'rho_gemm_m1.hip': ['-mllvm', '-amdgpu-mfma-vgpr-form']}


def _compile(src, verbose):
  obj = os.path.join(OBJ, os.path.splitext(src)[0] + '.o')
  cmd = [HIPCC] + FLAGS + EXTRA.get(src, []) + ['-c', os.path.join(CSRC, src), '-o', obj]
  r = subprocess.run(cmd, capture_output=True, text=True)
  out = '\n'.join(l for l in (r.stdout + r.stderr).splitlines()
                  if 'packed-fp32-ops' not in l)
  if verbose and out.strip():
    print(out)
  if r.returncode:
    raise RuntimeError('hipcc failed for %s:\n%s' % (src, out))
  return obj


def build(verbose=False):
  os.makedirs(OBJ, exist_ok=True)
  with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
    objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
  tmp = LIB + '.tmp'
  r = subprocess.run([HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs,
                     capture_output=True, text=True)
  if r.returncode:
    raise RuntimeError('link failed:\n' + r.stdout + r.stderr)
  os.replace(tmp, LIB)
  return LIB


if __name__ == '__main__':
  print(build(verbose='-v' in sys.argv))
