import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def _have_gpu() -> bool:
    try:
        from mrilab_amd import engine
        return engine.device_count() > 0
    except Exception:  # noqa: BLE001
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skips GPU tests when no device is visible (e.g. the CPU container)."""
    if not _have_gpu():
        pytest.skip("no HIP device visible")
    return True


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def adc_capture_slots(seq):
    """For each ADC sample of a scan, the RunMode-1 capture slot holding the
    state it samples: replays the decoder's bookkeeping (time points merged
    while ts repeats; the ADC of the step starting at uts[u] reads the state
    captured at slot u, DoScanAtCPU.cpp:445-491, 544-552)."""
    import numpy as np
    uts, ts, fl = seq["utsLine"], seq["tsLine"], seq["flagsLine"]
    slots = []
    utsi = 0
    i = 0
    adci = 0
    while i < ts.size - 1:
        f = fl[:, i].copy()
        j = i
        while j + 1 < ts.size and ts[j + 1] == ts[j]:
            f += fl[:, j + 1]
            j += 1
            if j == ts.size - 1:
                break
        i = j + 1
        cur = utsi
        utsi = min(utsi + 1, uts.size - 1)
        dtv = uts[utsi] - uts[cur]
        adc = 0
        if f[4] == 1:
            adc = seq["ADCLine"][adci]
            adci += 1
        if dtv > 0 and adc == 1:
            slots.append(cur)
    return np.asarray(slots, dtype=np.int64)
