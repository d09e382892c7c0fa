"""Builds tests/c_caller/c_caller (a plain C program linked against
qradient_amd/libqradient_amd.so) with gcc.  Called by __graft_entry__.build()
after the library; the binary is git-ignored and travels with the tree."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, '..', '..'))
SRC = os.path.join(HERE, 'c_caller.c')
OUT = os.path.join(HERE, 'c_caller')


def build():
    lib_dir = os.path.join(ROOT, 'qradient_amd')
    cmd = ['gcc', '-O2', '-std=c11', '-Wall', '-I', os.path.join(ROOT, 'include'), SRC, '-o', OUT + '.tmp',
           '-L', lib_dir, '-lqradient_amd', '-Wl,-rpath,$ORIGIN/../../qradient_amd',
           '-Wl,-rpath-link,/opt/rocm/lib']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('c_caller build failed:\n' + r.stderr)
    os.replace(OUT + '.tmp', OUT)
    return OUT


if __name__ == '__main__':
    print(build())
