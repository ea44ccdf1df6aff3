"""Build libmidfusion.so (HIP, gfx950) in-tree with hipcc.

Flags: -O3, gfx950 only, -ffp-contract=off (expressions round exactly as
written, so block keys match the reference bit for bit), packed-FP32
instruction selection disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'libmidfusion.so')
SOURCES = ['mf_pre.hip', 'mf_track.hip', 'mf_map.hip', 'mf_raycast.hip', 'mf_api.hip', 'mf_io.hip', 'mf_render.hip', 'mf_mesh.hip']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=gfx950', '-ffp-contract=off', '-fno-fast-math',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wno-unused-result']


def build(verbose=False):
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)

    def comp(src):
        obj = os.path.join(objdir, src + '.o')
        cmd = [HIPCC] + FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError('hipcc failed on %s:\n%s' % (src, r.stderr[-6000:]))
        if verbose and r.stderr.strip():
            print(r.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(comp, SOURCES))
    tmp = OUT + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=gfx950'] + objs + ['-o', tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n' + r.stderr[-4000:])
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
