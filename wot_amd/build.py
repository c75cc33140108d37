"""Build libwotk.so (all HIP sources under csrc/) for gfx950 with hipcc.

    python -m wot_amd.build          # compiles in place: wot_amd/_lib/libwotk.so
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIBDIR = os.path.join(HERE, '_lib')
LIB = os.path.join(LIBDIR, 'libwotk.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

DEVICE_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC',
    # keep the compiler from emitting packed FP32 VALU ops
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


# per-source device flags: the never-stored sweep keeps its MFMA accumulators
# in ordinary VGPRs (no AGPR copies before the per-entry exponentials)
SOURCE_FLAGS = {
    'fused.hip': ['-mllvm', '-amdgpu-mfma-vgpr-form'],
}


def sources():
    return sorted(glob.glob(os.path.join(CSRC, '*.hip')) + glob.glob(os.path.join(CSRC, '*.cpp')))


def build(verbose=False, jobs=8):
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    procs = []
    for src in sources():
        obj = os.path.join(LIBDIR, os.path.basename(src) + '.o')
        lang = ['-x', 'hip'] if src.endswith('.hip') else []
        cmd = [HIPCC] + DEVICE_FLAGS + SOURCE_FLAGS.get(os.path.basename(src), []) + lang + ['-c', src, '-o', obj]
        if src.endswith('.cpp'):
            cmd = [HIPCC, '-O3', '-std=c++17', '-fPIC', '-c', src, '-o', obj]
        objs.append(obj)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0), verbose)
    for p in procs:
        _wait(p, verbose)
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs
    _wait((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)), verbose)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def _wait(item, verbose):
    cmd, p = item
    out, _ = p.communicate()
    text = out.decode(errors='replace')
    if p.returncode != 0:
        raise RuntimeError('hipcc failed (%d): %s\n%s' % (p.returncode, ' '.join(cmd), text))
    if verbose and text.strip():
        print(text)


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
