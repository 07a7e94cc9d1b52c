"""CPU oracle for the modulated Pyramid-WFS hot path (numpy restatement of OOPAO).

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import anything from this package, and
only as the checker / the timed CPU baseline. The product package
``oopao_amd`` never imports it: its hot path runs on the HIP library and fails
loudly when that library is missing.

Every function cites the reference file:line it restates (paths are relative to
the OOPAO repository root). Which checks pin each function is listed in the
module docstrings; functions without such a check say "parity unpinned".
"""
