"""The oracle: CPU restatement of the reference's contracted ERI algorithm.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker or as the timed CPU baseline.  The product (``intception_amd``) never
imports it and has no CPU fallback.

Contents:
  eri_oracle.c / oracle.py  -- C restatement of the generated ``<full>_main``
                               (Boys, base, 2-layer VRR a->c->d, contraction,
                               HRR a->b, sph_trans, output copy) + ctypes wrapper
  mp_eri.py                 -- independent arbitrary-precision McMurchie-Davidson
                               ERI (mpmath) used to pin the C restatement
"""
