"""GPU blue-noise synthesis (ps_synthesise_blue_noise) against the oracle
restatement of generate_blue_noise(64) (crates/ps-clouds/src/noise.rs:
646-704) and the host generator: every rank byte identical; a render with the
synthesised map equals the render with the same map uploaded."""
import sys

import numpy as np
import pytest

import oracle.blue_noise as BN
from pedalsky_amd import abi, scene
from pedalsky_amd.renderer import Renderer

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def r():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU (no CPU fallback exists)"
    rr = Renderer(0)
    yield rr
    rr.close()


def test_ranks_bit_identical(r):
    _log("case: blue-noise synthesis 64x64 on the GPU vs the oracle and the host generator")
    r.synthesise_blue_noise()
    got = r.read_blue_noise().reshape(-1)
    want = BN.generate()
    bad = int(np.count_nonzero(got != want))
    _log(f"  rank bytes differing from the oracle: {bad}")
    assert bad == 0
    assert np.array_equal(got, scene.blue_noise())
    # a second synthesis into the same context gives the same map
    r.synthesise_blue_noise()
    assert np.array_equal(r.read_blue_noise().reshape(-1), want)


def test_synthesised_map_drives_the_jitter(r):
    import torch
    w = scene.workload_small()
    r.load_workload(w)
    r.bake_luts()
    _log("case: render 320x180 with synthesised vs uploaded blue noise")
    bufs = r.alloc_outputs(w.width, w.height, abi.PS_PIXEL_F32, ("composite",))
    r.synthesise_blue_noise()
    r.render(r.image_out(bufs))
    r.synchronize()
    a = bufs["composite"].cpu().numpy()
    r.upload_blue_noise(w.blue)
    r.render(r.image_out(bufs))
    r.synchronize()
    torch.cuda.synchronize()
    b = bufs["composite"].cpu().numpy()
    assert np.array_equal(w.blue.reshape(-1), BN.generate())
    assert np.array_equal(a, b)


def test_read_errors(r):
    lib = r.lib
    assert lib.ps_read_blue_noise(r.ctx, None) == abi.PS_ERR_INVALID_ARGUMENT
    assert lib.ps_synthesise_blue_noise(None) == abi.PS_ERR_INVALID_ARGUMENT
