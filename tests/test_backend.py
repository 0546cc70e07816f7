"""The backend boundary on the CPU: the job interface's error convention
(PythonBackendWrapper::ConvertException, Backend/Backend/PythonBackendWrapper.cpp:49-60,
and dplus.Backend.BackendError, PythonInterface/dplus/Backend.py:4-8), status
JSON shape and model metadata.  Nothing here launches a kernel."""
import json
import os

import pytest

from conftest import gpu_available

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _state(name="o_explicit_SS334"):
    rec = json.load(open(os.path.join(GOLDEN, name + ".json")))
    return {"state": rec["state"], "x": rec["q"], "y": rec["intensity"]}


def test_illegal_json_is_code_9():
    from dplus_mi355x.wrappers import BackendWrapper
    w = BackendWrapper()
    with pytest.raises(RuntimeError) as e:
        w.start_generate("{\"state\": [1, 2", True)
    d = json.loads(str(e.value))
    assert d["code"] == 9 and "JSON" in d["message"]


def test_backend_error_mirror():
    from dplus_mi355x.backend import Backend, BackendError
    b = Backend()
    with pytest.raises(BackendError) as e:
        b.start_generate("not json at all", True)
    assert e.value.error_code == 9


@pytest.mark.skipif(gpu_available(), reason="checks the behaviour without a GPU")
def test_no_gpu_is_code_31():
    from dplus_mi355x.backend import Backend, BackendError
    b = Backend()
    with pytest.raises(BackendError) as e:
        b.start_generate(json.dumps(_state()), True)
    assert e.value.error_code == 31
    with pytest.raises(BackendError) as e:
        b.check_capabilities(True)
    assert e.value.error_code == 31


def test_idle_status_and_no_results():
    from dplus_mi355x.backend import Backend, BackendError
    b = Backend()
    st = b.get_job_status()
    assert set(st) == {"isRunning", "progress", "code", "message"}
    assert st["isRunning"] is False
    with pytest.raises(BackendError):
        b.get_generate_results()
    with pytest.raises(BackendError) as e:
        b.get_amp(12345)
    assert e.value.error_code == 14  # ERROR_MODELNOTFOUND


def test_get_pdb_without_a_job_is_code_3():
    """GetPDB before any job: ERROR_JOBNOTFOUND (LocalBackend.cpp:973-975)."""
    from dplus_mi355x.backend import Backend, BackendError
    with pytest.raises(BackendError) as e:
        Backend().get_pdb(2)
    assert e.value.error_code == 3


def test_metadata_lists_hot_path_models():
    from dplus_mi355x.backend import Backend
    md = Backend().get_all_model_metadata()
    types = {m["Type"] for m in md["Models"]}
    assert {"PDB", "AMP", ",26", ",25"} <= types
    assert any("VEGAS" in m for m in md["OrientationMethods"])


def test_initialize_cache_and_stop_without_job():
    from dplus_mi355x.backend import Backend
    b = Backend()
    b.initialize_cache("/tmp")
    b.stop()
    assert b.get_model_ptrs() == []
