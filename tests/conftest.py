import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


GOLDEN_CASES = ("cfg0", "cfg0_k", "cfg0_fullmesh", "cfg0_pad", "cfg0_nonl", "ortho", "cfg1_b2")

# the specialised mesh lengths of the device kernels beyond 24 and 72 (the
# golden cases): (workload, kmscr(1:3) raised so that size_fft gives n, n)
ENLARGED_MESHES = ((0, 23, 48), (0, 31, 64), (1, 39, 80), (1, 47, 96), (1, 53, 108), (1, 63, 128))


def mesh_config(cfg, kmax, nbands):
    """Workload cfg with the FFT box enlarged to kmscr(1:3) = kmax (same
    sphere, potential and projectors; more mesh points)."""
    from cpw2000_amd import synth

    return synth.make_config(cfg, nbands=nbands, kmax_override=[kmax] * 3)


def load_golden(name):
    """(inputs dict, psi slice, reference hpsi) of a golden fixture; checks the
    regenerated inputs against the stored digests."""
    from cpw2000_amd import synth
    from tests.golden.make_golden import digest

    z = np.load(os.path.join(GOLDEN, f"hkpsi_{name}.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    kw = dict(meta["kwargs"])
    for k in ("kpt", "kmax_override"):
        if k in kw:
            kw[k] = tuple(kw[k])
    d = synth.make_config(**kw)
    dg = digest(d)
    for k, v in meta["digest"].items():
        np.testing.assert_allclose(dg[k], v, rtol=1e-12, atol=1e-12,
                                   err_msg=f"synthetic generator drifted ({name}: {k})")
    if not meta["projectors"]:
        d = dict(d)
        d["nanl"] = 0
    psi = np.asfortranarray(d["psi"][:, : meta["nbands"]])
    return d, psi, z["hpsi"]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O

    O.orc_lib()
    return O
