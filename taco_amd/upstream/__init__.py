"""taco itself with this package's GPU backend plugged in.

libtaco_upstream.so is taco (the reference's own sources, compiled where
they lie by taco_amd/upstream/Makefile) with CodeGen_HIP (codegen_hip.cpp),
the HIP Module (module_hip.cpp) and the cuda.h runtime forwarders
(cuda_hip.cpp) in place of its CUDA backend.  libtaco_dropin.so runs the
reference's GPU test statements through it (drop_in_tests.cpp).

This module loads the latter; there is no fallback: if the libraries were
not built, every call raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
DROPIN = os.path.join(_HERE, "libtaco_dropin.so")
UPSTREAM = os.path.join(_HERE, "libtaco_upstream.so")

_lib = None


def available():
    return os.path.exists(DROPIN) and os.path.exists(UPSTREAM)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"{DROPIN} missing: build it with make -C taco_amd/upstream")
        from .. import _lib as hip
        hip.lib()  # libtaco_hip.so first, as the generated modules expect
        _lib = ctypes.CDLL(DROPIN)
        _lib.taco_dropin_list.restype = ctypes.c_char_p
        _lib.taco_dropin_run.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                         ctypes.c_int]
        _lib.taco_dropin_emit.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    return _lib


def names():
    """The statements, in the order the tests run them."""
    return lib().taco_dropin_list().decode().split(",")


def run(name, compile_only=False):
    """(ok, message): compile, assemble, compute and check one statement
    (compile_only: lower and compile its modules, no GPU needed)."""
    msg = ctypes.create_string_buffer(4096)
    rc = lib().taco_dropin_run(name.encode(), 1 if compile_only else 0, msg, len(msg))
    return rc == 0, msg.value.decode(errors="replace")


def emit(name):
    """The HIP module CodeGen_HIP prints for a statement (no compilation)."""
    buf = ctypes.create_string_buffer(1 << 21)
    rc = lib().taco_dropin_emit(name.encode(), buf, len(buf))
    text = buf.value.decode(errors="replace")
    if rc != 0:
        raise RuntimeError(f"{name}: {text}")
    return text


def prebuild():
    """Compiles every statement's modules into the module cache (run in a
    process of its own: taco's runtime switches the library to unified
    memory)."""
    out = {}
    for n in names():
        out[n] = run(n, compile_only=True)
    return out

