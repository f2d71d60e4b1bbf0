"""CPU oracle for the table1 NUFFT hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``mrrt/nufft_amd`` may import, call or
link anything under ``oracle/``.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker
(or as the timed CPU baseline), never as the thing measured or shipped.

Contents
--------
nufft_oracle
    numpy/scipy restatement of the reference's table1 NUFFT
    (mrrt/nufft/_nufft.py, _kaiser_bessel.py, _kernels.py, _utils.py,
    _dtft.py and the C kernels in _extensions/c/nufft_table.template.c).
ref_ctypes
    ctypes bindings for ``oracle/_ref/libnufft_table_ref.so``: the reference's
    own C kernels compiled from /root/reference by ``oracle/build_ref.sh``.

Pinning status of each function is stated in the module docstrings.
"""
