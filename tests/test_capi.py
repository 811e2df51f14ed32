"""The C ABI library: loads, exports every symbol include/pa_admm.h declares,
fills the reference's config defaults, and reports errors as status codes
(no GPU calls here).  CPU only."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "pa_admm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pa_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(pa):
    L = pa.lib()
    names = _declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(pa.EXPORTS)


def test_config_defaults_follow_reference(pa):
    c = pa.AdmmConfigC()
    pa.lib().pa_admm_config_default(C.byref(c))
    # ADMMSolverConfig defaults (admm_config.h:20-53)
    assert (c.admm_max_iter, c.dcd_interval, c.gs_max_iter) == (25, 5, 10)
    assert (c.warmstart_ue, c.warmstart_uc, c.enable_contact) == (1, 1, 1)
    assert (c.mu, c.kappa, c.beta, c.damp_kl, c.damp_km) == (0.5, 1e6, 1e5, 1e-4, 0.0)
    assert c.g == pytest.approx(0.98)
    d = pa.AdmmConfig()
    for f, _ in pa.AdmmConfigC._fields_:
        assert getattr(c, f) == pytest.approx(int(getattr(d, f)) if isinstance(getattr(d, f), bool) else getattr(d, f)), f


def test_null_and_order_errors(pa):
    L = pa.lib()
    assert L.pa_scene_generate(9, 0, 0, C.c_double(0), C.c_double(0), C.byref(C.c_void_p())) == pa.PA_ERR_INVALID
    assert L.pa_scene_get_info(None, None) == pa.PA_ERR_INVALID
    assert L.pa_backend_step(None, None) == pa.PA_ERR_INVALID
    assert L.pa_detector_detect(None, None, None) == pa.PA_ERR_INVALID
    assert b"null" in L.pa_last_error_message()


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a GPU is visible")
def test_backend_without_gpu_reports_no_device(pa):
    h = C.c_void_p()
    o = pa.DeviceOptsC(0, 0)
    assert pa.lib().pa_backend_create(C.byref(o), C.byref(h)) == pa.PA_ERR_NO_DEVICE
    assert not h.value


def test_oracle_library_loads():
    from oracle import pyoracle

    L = pyoracle.lib()
    for n in ("orc_create", "orc_step", "orc_detect", "orc_dist_pt", "orc_llt_solve"):
        assert hasattr(L, n)
