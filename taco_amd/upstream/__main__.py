"""python -m taco_amd.upstream: compile every drop-in statement's modules
into the module cache (no GPU needed)."""
import sys

from . import prebuild

failed = 0
for name, (ok, msg) in prebuild().items():
    print(f"{name}: {'ok' if ok else 'FAILED'} {msg}", flush=True)
    failed += not ok
sys.exit(1 if failed else 0)
