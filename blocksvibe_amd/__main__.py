"""python -m blocksvibe_amd [blocks_3d options] | batch ...  (one BLAS/OpenMP thread per process)"""
import os
import sys

for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ.setdefault(_v, "1")

from .cli import main  # noqa: E402

sys.exit(main())
