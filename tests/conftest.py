import json
import subprocess
import sys
import threading
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

_spawner = None
_spawner_lock = threading.Lock()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def pytest_sessionstart(session):
    # the child-process launcher starts here, before any test can initialise
    # the GPU in this process (tests/spawner.py says why)
    global _spawner
    _spawner = subprocess.Popen([sys.executable, "-u", str(Path(__file__).resolve().parent / "spawner.py")],
                                stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True, cwd=str(ROOT))


def pytest_sessionfinish(session, exitstatus):
    global _spawner
    if _spawner is not None:
        try:
            _spawner.stdin.close()
            _spawner.wait(timeout=30)
        except Exception:
            _spawner.kill()
        _spawner = None


@pytest.fixture(scope="session")
def spawn():
    """run(argv, cwd=None, timeout=120) -> (returncode, stdout, stderr),
    executed by the session's launcher process, not by this one."""
    def run(argv, cwd=None, timeout=120):
        if _spawner is None or _spawner.poll() is not None:
            raise RuntimeError("the child-process launcher is not running")
        with _spawner_lock:
            _spawner.stdin.write(json.dumps({"argv": [str(a) for a in argv], "cwd": None if cwd is None else str(cwd),
                                             "timeout": timeout}) + "\n")
            _spawner.stdin.flush()
            line = _spawner.stdout.readline()
        if not line:
            raise RuntimeError("the child-process launcher ended")
        out = json.loads(line)
        return out["returncode"], out["stdout"], out["stderr"]
    return run


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_lib():
    from parallellocallyoptimalapproximatesparsematrixinverses_amd import _lib
    _lib.load()
    return _lib
