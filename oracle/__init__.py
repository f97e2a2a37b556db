"""FP64 CPU oracle for the smuthi coupling-solve and near-field hot path.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import anything from this package, and
only as the checker or the timed CPU baseline. The product package
``smuthi_amd`` never imports it and has no CPU fallback.

Each function restates the algorithm of the reference file:line it cites
(paths relative to the smuthi source tree). Every module header lists how its
functions are pinned: golden literals of the reference's own tests, closed
forms, identities, or brute force on tiny inputs. A function pinned by none of
them says "parity unpinned" in its docstring.
"""
