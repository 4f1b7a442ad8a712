"""Child-process launcher for the test session.

The test process starts this helper once, before any test touches the GPU,
and GPU tests that need a child process ask it to run one.  A GPU test process
therefore never spawns: CPython 3.10's subprocess (with cwd= or close_fds,
its defaults) creates the child with vfork(), so the child runs in the
parent's address space until execve while the parent's HIP runtime threads
keep running; preloaded libraries interpose on that path too.  This helper
never loads HIP, so its own vfork is of a single-threaded process.

Protocol: one JSON object per line on stdin, {"argv": [...], "cwd": str|null,
"timeout": seconds}; one JSON object per line on stdout, {"returncode": int,
"stdout": str, "stderr": str} (returncode -9 on a timeout).
"""
import json
import subprocess
import sys


def main():
    for line in sys.stdin:
        req = json.loads(line)
        try:
            p = subprocess.run(req["argv"], cwd=req.get("cwd"), capture_output=True, text=True,
                               timeout=req.get("timeout", 120))
            out = {"returncode": p.returncode, "stdout": p.stdout, "stderr": p.stderr}
        except subprocess.TimeoutExpired as e:
            out = {"returncode": -9, "stdout": e.stdout or "", "stderr": (e.stderr or "") + "\n[timeout]"}
        except OSError as e:
            out = {"returncode": -1, "stdout": "", "stderr": str(e)}
        sys.stdout.write(json.dumps(out) + "\n")
        sys.stdout.flush()


if __name__ == "__main__":
    main()
