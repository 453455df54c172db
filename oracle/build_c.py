"""Build the oracle's C helpers (test infrastructure) with gcc into oracle/_build/."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, '_build', 'libpolesum.so')


def build() -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    subprocess.run(['gcc', '-O3', '-march=x86-64-v2', '-fopenmp', '-fPIC', '-shared', '-std=c99',
                    os.path.join(HERE, 'c', 'pole_sum.c'), '-o', OUT + '.tmp', '-lm'], check=True)
    os.replace(OUT + '.tmp', OUT)
    return OUT


if __name__ == '__main__':
    print(build())
