"""GPU: joins from several host threads at once.  Each thread owns a context (its own
stream and workspace) and also calls the drop-in PRO/NPO (one shared default
context, serialised by its lock); every result equals the oracle's.  The threads
release the GIL inside the ctypes calls, so the joins overlap on the device."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_concurrent_contexts_and_dropins(aj, oracle):
    cases = []
    for i in range(4):
        n = (1 << 16) << i  # 64K .. 512K: different geometries per thread
        R = aj.gen_unique(n, 100 + i)
        S = aj.gen_uniform(n + 13 * i, 1, n, 200 + i)
        cases.append((R, S, oracle.aggregate(R, S)))
    errors = []

    def work(i):
        try:
            R, S, want = cases[i]
            algo = aj.AJ_ALGO_PRO if i % 2 == 0 else aj.AJ_ALGO_NPO
            dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
            try:
                with aj.Context(0) as ctx:
                    for rep in range(5):
                        out = ctx.join(dR.ptr, len(R), dS.ptr, len(S), algo)
                        if (out.count, out.checksum) != want:
                            errors.append(("ctx", i, rep, out.count, out.checksum, want))
                        got = (aj.PRO if rep % 2 == 0 else aj.NPO)(R, S, 4)
                        if (got.count, got.checksum) != want:
                            errors.append(("dropin", i, rep, got.count, got.checksum, want))
            finally:
                dR.free()
                dS.free()
        except Exception as e:  # reported below, in the main thread
            errors.append(("exception", i, repr(e)))

    print("case 4 threads x (own context + drop-in) x 5 joins", flush=True)
    threads = [threading.Thread(target=work, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors[:4]
