"""``Backend`` and ``BackendError`` of the reference's Python interface
(PythonInterface/dplus/Backend.py:1-112) over this library's BackendWrapper,
plus the embedded runner's generate loop (CalculationRunner.py:707-838)."""
import json

from .wrappers import BackendWrapper


class BackendError(RuntimeError):
    """Backend.py:4-8: error_code and message parsed from the wrapper's JSON text."""

    def __init__(self, re):
        d = json.loads(str(re))
        self.error_code = d["code"]
        super().__init__(d["message"])


def _guard(f):
    def g(*a, **k):
        try:
            return f(*a, **k)
        except RuntimeError as re:
            if isinstance(re, BackendError):
                raise
            raise BackendError(re) from None
    g.__name__, g.__doc__ = f.__name__, f.__doc__
    return g


class Backend:
    """Backend.py:10-112."""

    def __init__(self):
        self._wrapper = BackendWrapper()

    @_guard
    def check_capabilities(self, check_tdr=True):
        return self._wrapper.check_capabilities(check_tdr)

    @_guard
    def get_all_model_metadata(self):
        return json.loads(self._wrapper.get_all_model_metadata())

    @_guard
    def initialize_cache(self, cache_dir):
        self._wrapper.initialize_cache(cache_dir)

    @_guard
    def start_generate(self, data, useGPU=True):
        self._wrapper.start_generate(data.encode("utf-8") if isinstance(data, str) else data, useGPU)

    @_guard
    def start_generate_2D(self, data, useGPU=True):
        self._wrapper.start_generate_2D(data.encode("utf-8") if isinstance(data, str) else data, useGPU)

    @_guard
    def get_generate_2D_results(self):
        return json.loads(self._wrapper.get_generate_2D_results())

    @_guard
    def get_job_status(self):
        return json.loads(self._wrapper.get_job_status())

    @_guard
    def get_generate_results(self):
        return json.loads(self._wrapper.get_generate_results())

    @_guard
    def save_amp(self, modelptr, path):
        self._wrapper.save_amp(modelptr, path)

    @_guard
    def get_amp(self, modelptr):
        return self._wrapper.get_amp(modelptr)

    @_guard
    def get_pdb(self, modelptr):
        return self._wrapper.get_pdb(modelptr)

    @_guard
    def get_model_ptrs(self):
        return self._wrapper.get_model_ptrs()

    @_guard
    def stop(self):
        self._wrapper.stop()


class EmbeddedLocalRunner:
    """The in-process runner of CalculationRunner.py:707-838 for generate jobs:
    ``generate(args)`` takes the CalculationInput args dict ({"state", "x", ...}),
    waits for the job and returns the results dict ({"Headers", "Graph"})."""

    def __init__(self):
        self.wrapper = Backend()

    def check_capabilities(self, check_tdr=True):
        self.wrapper.check_capabilities(check_tdr)

    def generate_async(self, args, use_gpu=True):
        self.wrapper.start_generate(json.dumps(args), useGPU=use_gpu)

    def wait(self, poll=0.01):
        import time
        status = self.wrapper.get_job_status()
        while status and status["isRunning"] and status["code"] == -1:
            time.sleep(poll)
            status = self.wrapper.get_job_status()
        return status

    def generate(self, args, use_gpu=True):
        self.generate_async(args, use_gpu)
        status = self.wait()
        if status["code"] not in (0,):
            raise BackendError(RuntimeError(json.dumps({"code": status["code"], "message": status["message"]})))
        return self.wrapper.get_generate_results()

    def generate2D(self, args, use_gpu=True):
        """CalculationRunner.py:741-756: the 2-D job; returns {"Headers", "2DGraph"}."""
        self.wrapper.start_generate_2D(json.dumps(args), useGPU=use_gpu)
        status = self.wait()
        if status["code"] not in (0,):
            raise BackendError(RuntimeError(json.dumps({"code": status["code"], "message": status["message"]})))
        return self.wrapper.get_generate_2D_results()

    def get_job_status(self):
        return self.wrapper.get_job_status()

    def save_amp(self, modelptr, path):
        self.wrapper.save_amp(modelptr, path)

    def get_amp(self, modelptr):
        return self.wrapper.get_amp(modelptr)

    def get_pdb(self, model_ptr):
        return self.wrapper.get_pdb(model_ptr)

    def get_model_ptrs(self):
        return self.wrapper.get_model_ptrs()

    def stop_generate(self):
        self.wrapper.stop()
