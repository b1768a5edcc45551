"""Test-infrastructure oracle: an FP64 NumPy restatement (``tgv_oracle``) and
an FP32 C/OpenMP restatement (``cref``) of the reference's primal-dual TGV
path.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package; the product never does."""
