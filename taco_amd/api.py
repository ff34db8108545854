"""taco's tensor API for the hot path, over libtaco_hip.

The names and call order follow taco's C++ API (reference
include/taco/tensor.h, include/taco/index_notation/index_notation.h,
src/tensor.cpp:605-858): a ``Tensor`` has a name, dimensions, a ``Format``
and a datatype; ``y[i] = A[i, j] * x[j]`` records the assignment (taco
writes ``y(i) = A(i,j) * x(j)``); ``get_assignment().concretize()`` gives
an ``IndexStmt`` that takes taco's scheduling commands (``split``, ``pos``,
``fuse``, ``reorder``, ``bound``, ``precompute``, ``unroll``,
``parallelize``, reference index_notation.cpp:1819-2037) and
``compile`` / ``assemble`` / ``compute`` / ``evaluate`` run it.

The statements of the hot path are supported — SpMV ``y(i) = A(i,j) *
x(j)``, SpMM ``C(i,k) = A(i,j) * B(j,k)`` with dense B and C, and SpGEMM
with all three CSR — under the GPU schedules of the reference's scheduling
tests (test/tests-scheduling-eval.cpp:193-268) and taco's default GPU
schedule (parallelizeOuterLoop, transformations.cpp:1562-1606).  The
recognised statement and schedule become a ``taco_hip_gen_spec``; CodeGen_HIP
emits the module and the JIT loader compiles it.  Anything else raises
``TacoError`` (taco raises ``TacoException`` for unsupported schedules).
"""
import numpy as np

from . import _lib
from . import codegen
from . import descriptors as D


class TacoError(RuntimeError):
    pass


# ---------------------------------------------------------------- types --

class Datatype:
    def __init__(self, name, np_dtype, csize, tag):
        self.name, self.np, self.csize, self.tag = name, np_dtype, csize, tag

    def __repr__(self):
        return self.name


Float64 = Datatype("double", np.float64, 64, "f64")
Float32 = Datatype("float", np.float32, 32, "f32")


class ModeFormat:
    def __init__(self, name):
        self.name = name

    def __repr__(self):
        return self.name


Dense = ModeFormat("dense")
Sparse = ModeFormat("compressed")
Compressed = Sparse
dense, sparse, compressed = Dense, Sparse, Sparse


class Format:
    def __init__(self, modes):
        self.modes = list(modes) if isinstance(modes, (list, tuple)) else [modes]

    def __eq__(self, other):
        return isinstance(other, Format) and [m.name for m in self.modes] == [m.name for m in other.modes]

    def __repr__(self):
        return "(" + ",".join(m.name for m in self.modes) + ")"


CSR = Format([Dense, Sparse])
dense_vector = Format([Dense])
dense_matrix = Format([Dense, Dense])


class ParallelUnit:
    NotParallel, DefaultUnit, GPUBlock, GPUWarp, GPUThread, CPUThread, CPUVector = (
        "NotParallel", "DefaultUnit", "GPUBlock", "GPUWarp", "GPUThread", "CPUThread",
        "CPUVector")


class OutputRaceStrategy:
    IgnoreRaces, NoRaces, Atomics, Temporary, ParallelReduction = (
        "IgnoreRaces", "NoRaces", "Atomics", "Temporary", "ParallelReduction")


class BoundType:
    MinExact, MinConstraint, MaxExact, MaxConstraint = (
        "MinExact", "MinConstraint", "MaxExact", "MaxConstraint")


# ------------------------------------------------------- index notation --

class IndexVar:
    _count = 0

    def __init__(self, name=None):
        IndexVar._count += 1
        self.name = name or f"i{IndexVar._count}"

    def __repr__(self):
        return self.name


class IndexExpr:
    def __mul__(self, other):
        return Mul(self, other)


class Access(IndexExpr):
    def __init__(self, tensor, ivars):
        self.tensor, self.vars = tensor, list(ivars)

    def __repr__(self):
        return f"{self.tensor.name}({','.join(map(repr, self.vars))})"


class Mul(IndexExpr):
    def __init__(self, a, b):
        self.a, self.b = a, b

    def __repr__(self):
        return f"{self.a!r} * {self.b!r}"


class Assignment:
    def __init__(self, lhs, rhs):
        self.lhs, self.rhs = lhs, rhs

    def concretize(self):
        return IndexStmt(self, [])

    def __repr__(self):
        return f"{self.lhs!r} = {self.rhs!r}"


class TensorVar:
    """A temporary of a schedule (precompute); recorded, not materialised."""

    def __init__(self, name, dtype=Float64, dims=None, fmt=dense):
        self.name, self.dtype, self.dims, self.format = name, dtype, dims, fmt


class IndexStmt:
    """A concrete statement plus the scheduling commands applied to it.

    Each command returns a new statement, as in taco."""

    def __init__(self, assignment, commands):
        self.assignment, self.commands = assignment, list(commands)

    def _with(self, *cmd):
        return IndexStmt(self.assignment, self.commands + [cmd])

    def split(self, i, outer, inner, factor):
        return self._with("split", i, outer, inner, int(factor))

    def pos(self, i, ipos, access):
        return self._with("pos", i, ipos, access)

    def fuse(self, i, j, f):
        return self._with("fuse", i, j, f)

    def reorder(self, order):
        return self._with("reorder", list(order))

    def bound(self, i, bounded, value, bound_type):
        return self._with("bound", i, bounded, int(value), bound_type)

    def precompute(self, expr, i, iw, workspace):
        return self._with("precompute", expr, i, iw, workspace)

    def unroll(self, i, factor):
        return self._with("unroll", i, int(factor))

    def parallelize(self, i, unit, strategy):
        return self._with("parallelize", i, unit, strategy)

    def __repr__(self):
        return f"{self.assignment!r} with {[c[0] for c in self.commands]}"


# ------------------------------------------------------ recognising work --

def _statement_kind(asg):
    """('spmv' | 'spmm' | 'spgemm', A access, other operand access)."""
    lhs, rhs = asg.lhs, asg.rhs
    if not isinstance(rhs, Mul) or not isinstance(rhs.a, Access) or not isinstance(rhs.b, Access):
        raise TacoError(f"unsupported statement {asg!r}: the HIP backend runs "
                        "y(i)=A(i,j)*x(j), C(i,k)=A(i,j)*B(j,k)")
    a, b = rhs.a, rhs.b
    if b.tensor.format == CSR and a.tensor.format != CSR:
        a, b = b, a
    if a.tensor.format != CSR or len(a.vars) != 2:
        raise TacoError(f"unsupported statement {asg!r}: the sparse operand must be CSR")
    i, j = a.vars
    if len(lhs.vars) == 1 and len(b.vars) == 1:
        if lhs.vars[0] is not i or b.vars[0] is not j or b.tensor.format != dense_vector:
            raise TacoError(f"unsupported statement {asg!r}")
        return "spmv", a, b
    if len(lhs.vars) == 2 and len(b.vars) == 2:
        k = lhs.vars[1]
        if lhs.vars[0] is not i or b.vars[0] is not j or b.vars[1] is not k:
            raise TacoError(f"unsupported statement {asg!r}")
        if b.tensor.format == dense_matrix and lhs.tensor.format == dense_matrix:
            return "spmm", a, b
        if b.tensor.format == CSR and lhs.tensor.format == CSR:
            return "spgemm", a, b
    raise TacoError(f"unsupported statement {asg!r}")


def _ops(cmds):
    return [c[0] for c in cmds if c[0] not in ("reorder", "precompute", "unroll")]


def _units(cmds):
    return {c[1]: (c[2], c[3]) for c in cmds if c[0] == "parallelize"}


def _recognise(stmt, kind, a):
    """The taco_hip_gen_spec fields (family, block, nnz/thread, warp, rows,
    bound) of the statement's GPU schedule."""
    cmds = stmt.commands
    i, j = a.vars
    ops = _ops(cmds)
    units = _units(cmds)
    if kind == "spgemm":
        if any(u[0] in (ParallelUnit.GPUWarp,) for u in units.values()):
            raise TacoError("SpGEMM: only the two-phase row-parallel schedule is supported")
        return (_lib.GEN_TWO_PHASE, 256, 0, 0, 0, 0)
    if not cmds:  # taco's default GPU schedule: split(i, 256)
        return (_lib.GEN_DEFAULT, 256, 0, 0, 0, 0)
    sp = [c for c in cmds if c[0] == "split"]
    if ops == ["split", "parallelize", "parallelize"]:
        _, v, outer, inner, f = sp[0]
        if (v is i and units.get(outer, (None,))[0] == ParallelUnit.GPUBlock
                and units.get(inner, (None,))[0] == ParallelUnit.GPUThread):
            return (_lib.GEN_DEFAULT, f, 0, 0, 0, 0)
    if kind == "spmv" and ops == ["fuse", "pos", "split", "split", "split", "parallelize",
                                  "parallelize", "parallelize"]:
        fuse = next(c for c in cmds if c[0] == "fuse")
        pos = next(c for c in cmds if c[0] == "pos")
        s1, s2, s3 = sp
        ok = (fuse[1] is i and fuse[2] is j and pos[1] is fuse[3] and s1[1] is pos[2]
              and s2[1] is s1[3] and s3[1] is s2[3]
              and units.get(s1[2], (None,))[0] == ParallelUnit.GPUBlock
              and units.get(s2[2], (None,))[0] == ParallelUnit.GPUWarp
              and units.get(s3[2], (None, None)) == (ParallelUnit.GPUThread,
                                                    OutputRaceStrategy.Atomics))
        if ok:
            per_thread = s3[4]
            block = s1[4] // per_thread
            warp = s2[4] // per_thread
            return (_lib.GEN_NNZ_SPLIT, block, per_thread, warp, 0, 0)
    if kind == "spmv" and ops == ["fuse", "split", "split", "parallelize", "parallelize"]:
        # fuse(i,j,f) split over rows and entries together, without pos():
        # merge-path load balancing (K merge items per thread).
        fuse = next(c for c in cmds if c[0] == "fuse")
        s1, s2 = sp
        ok = (fuse[1] is i and fuse[2] is j and s1[1] is fuse[3] and s2[1] is s1[3]
              and units.get(s1[2], (None,))[0] == ParallelUnit.GPUBlock
              and units.get(s2[2], (None, None)) == (ParallelUnit.GPUThread,
                                                    OutputRaceStrategy.Atomics))
        if ok:
            per_thread = s2[4]
            return (_lib.GEN_MERGE, s1[4] // per_thread, per_thread, 0, 0, 0)
    if kind == "spmv" and ops == ["pos", "split", "parallelize", "parallelize"]:
        # one workgroup per row; the thread Temporary without a GPUWarp
        # parent reduces with GPUBlockReduction (transformations.cpp:1049-1056)
        pos = next(c for c in cmds if c[0] == "pos")
        s1 = sp[0]
        ok = (pos[1] is j and s1[1] is pos[2]
              and units.get(i, (None,))[0] == ParallelUnit.GPUBlock
              and units.get(s1[3], (None, None)) == (ParallelUnit.GPUThread,
                                                    OutputRaceStrategy.Temporary))
        if ok:
            return (_lib.GEN_BLOCK_ROWS, s1[4], 0, 0, 0, 0)
    if kind == "spmv" and ops == ["split", "split", "pos", "split", "parallelize",
                                  "parallelize", "parallelize"]:
        pos = next(c for c in cmds if c[0] == "pos")
        s1, s2, s3 = sp
        ok = (s1[1] is i and s2[1] is s1[3] and pos[1] is j and s3[1] is pos[2]
              and units.get(s1[2], (None,))[0] == ParallelUnit.GPUBlock
              and units.get(s2[3], (None,))[0] == ParallelUnit.GPUWarp
              and units.get(s3[3], (None, None)) == (ParallelUnit.GPUThread,
                                                    OutputRaceStrategy.Temporary))
        if ok:
            warp = s3[4]
            block = s2[4] * warp
            rows = s1[4] // block
            return (_lib.GEN_WARP_ROWS, block, 0, warp, max(rows, 1), 0)
    if kind == "spmm" and ops == ["split", "split", "split", "parallelize", "parallelize",
                                  "parallelize"]:
        # scheduleSpMMRowsGPU: rows split over blocks and warps (R rows per
        # warp), k split into 32-column chunks over the warp's threads.
        s1, s2, s3 = sp
        k = next(v for v in stmt.assignment.lhs.vars[1:])
        ok = (s1[1] is i and s2[1] is s1[3] and s3[1] is k
              and units.get(s1[2], (None,))[0] == ParallelUnit.GPUBlock
              and units.get(s2[2], (None,))[0] == ParallelUnit.GPUWarp
              and units.get(s3[2], (None, None)) == (ParallelUnit.GPUThread,
                                                    OutputRaceStrategy.Atomics))
        if ok and s1[4] % s2[4] == 0:
            rows = s2[4]
            block = (s1[4] // rows) * 32  # WARP_SIZE threads per warp
            return (_lib.GEN_WARP_ROWS, block, 0, s3[4], rows, 0)
    if kind == "spmm" and ops == ["fuse", "pos", "split", "split", "split", "bound",
                                  "parallelize", "parallelize", "parallelize"]:
        fuse = next(c for c in cmds if c[0] == "fuse")
        pos = next(c for c in cmds if c[0] == "pos")
        bnd = next(c for c in cmds if c[0] == "bound")
        s1, s2, s3 = sp
        ok = (fuse[1] is i and fuse[2] is j and pos[1] is fuse[3] and s1[1] is pos[2]
              and s2[1] is s1[3] and bnd[1] is s3[2]
              and units.get(s1[2], (None,))[0] == ParallelUnit.GPUBlock
              and units.get(s2[2], (None,))[0] == ParallelUnit.GPUWarp
              and units.get(s3[3], (None, None)) == (ParallelUnit.GPUThread,
                                                    OutputRaceStrategy.Atomics))
        if ok:
            per_warp = s2[4]
            warp = s3[4]
            block = (s1[4] // per_warp) * warp
            return (_lib.GEN_NNZ_SPLIT, block, per_warp, warp, 0, bnd[3])
    if kind == "spmm" and ops == ["fuse", "split", "split", "split", "bound",
                                  "parallelize", "parallelize", "parallelize"]:
        # scheduleSpMMGPU's form without pos(): the fused (i,j) space split
        # over rows and entries together, merge-path teams along k.
        fuse = next(c for c in cmds if c[0] == "fuse")
        bnd = next(c for c in cmds if c[0] == "bound")
        s1, s2, s3 = sp
        ok = (fuse[1] is i and fuse[2] is j and s1[1] is fuse[3] and s2[1] is s1[3]
              and bnd[1] is s3[2]
              and units.get(s1[2], (None,))[0] == ParallelUnit.GPUBlock
              and units.get(s2[2], (None,))[0] == ParallelUnit.GPUWarp
              and units.get(s3[3], (None, None)) == (ParallelUnit.GPUThread,
                                                    OutputRaceStrategy.Atomics))
        if ok:
            per_team = s2[4]
            warp = s3[4]
            block = (s1[4] // per_team) * warp
            return (_lib.GEN_MERGE, block, per_team, warp, 0, bnd[3])
    raise TacoError(f"schedule not supported by the HIP backend: {ops}")


# ------------------------------------------------------ backend toggles --
# pytaco exposes taco's GPU backend switch (reference python_bindings/src/
# pytaco.cpp:14, should_use_cuda_codegen; include/taco/cuda.h:15-28); the
# facade forwards the same switch to the HIP runtime.  The facade compiles
# through CodeGen_HIP only: with code generation off, compile() raises.

def should_use_cuda_codegen():
    """taco's should_use_CUDA_codegen(): whether statements compile to GPU
    code (here CodeGen_HIP)."""
    return bool(_lib.lib().taco_hip_should_use_codegen())


def set_cuda_codegen_enabled(enabled):
    """taco's set_CUDA_codegen_enabled()."""
    _lib.lib().taco_hip_set_codegen_enabled(1 if enabled else 0)


should_use_hip_codegen = should_use_cuda_codegen
set_hip_codegen_enabled = set_cuda_codegen_enabled


# ----------------------------------------------------- learned schedules --

_learned = {"enabled": False, "selector": None}


def set_learned_schedule_enabled(enabled):
    """With it on, compile() of an SpMV or SpMM statement that carries no
    scheduling commands — where taco applies its default GPU schedule — lets
    the learned selector (taco_amd.selector) pick the schedule at the first
    compute() on each operand matrix: one of the hand-written schedules, or
    for f64 SpMV the slab plan, which is built once per matrix, kept on the
    operand tensor and dropped when that tensor's values change.  The choice
    changes which correct kernel runs, never the result beyond the value
    tolerance.  Off by default: taco's default schedule is the default."""
    _learned["enabled"] = bool(enabled)


def is_learned_schedule_enabled():
    return _learned["enabled"]


class _Learned:
    """The compute() of a statement scheduled by the learned selector."""

    def __init__(self, kind, A):
        self.op = 0 if kind == "spmv" else 1
        self.A = A

    def _choose(self, a_desc, K):
        from . import autotune, selector
        if _learned["selector"] is None:
            _learned["selector"] = selector.Selector()
        sel = _learned["selector"]
        if selector.mfma_rule(a_desc, self.op, K):
            return selector.RULE_MFMA, selector.class_schedule(selector.RULE_MFMA, K)
        z = sel.scores(selector.features(a_desc, self.op, K))[0]
        lo, hi = selector.class_range(self.op)
        for c in lo + np.argsort(-z[lo:hi], kind="stable"):
            s = selector.class_schedule(int(c), K)
            if s is None:  # the slab plan: f64 only, within its limits
                plan, _ = autotune.slab_plan(a_desc)
                if plan is None:
                    continue
                self.A._plan = plan
            return int(c), s
        raise TacoError("learned schedule: no applicable class")

    def assemble(self, *descs):
        pass

    def compute(self, out_desc, a_desc, b_desc):
        K = 0 if self.op == 0 else int(out_desc.t.dimensions[1])
        choice = self.A._schedule_choice.get((self.op, K))
        if choice is None:
            choice = self._choose(a_desc, K)
            self.A._schedule_choice[(self.op, K)] = choice
        c, s = choice
        L = _lib.lib()
        if s is None:
            if self.A._plan is None:
                from . import autotune
                self.A._plan, _ = autotune.slab_plan(a_desc)
            self.A._plan.compute(out_desc, b_desc)
        elif self.op == 0:
            _lib.check(L.taco_hip_spmv_compute(out_desc.ptr, a_desc.ptr, b_desc.ptr, s),
                       "learned SpMV")
        else:
            _lib.check(L.taco_hip_spmm_compute(out_desc.ptr, a_desc.ptr, b_desc.ptr, s),
                       "learned SpMM")

    def chosen(self, K=0):
        """Name of the class picked for the operand (None before compute)."""
        from . import selector
        ch = self.A._schedule_choice.get((self.op, K))
        return None if ch is None else selector.class_name(ch[0])


# --------------------------------------------------------------- tensors --

class Tensor:
    """A named tensor with host values and, once used, device buffers."""

    def __init__(self, name, dims, fmt=None, dtype=Float64):
        self.name = name
        self.dims = [int(d) for d in dims]
        self.format = fmt if fmt is not None else Format([Dense] * len(self.dims))
        if len(self.format.modes) != len(self.dims):
            raise TacoError(f"{name}: format {self.format!r} does not match order {len(self.dims)}")
        if self.format not in (CSR, dense_vector, dense_matrix):
            raise TacoError(f"{name}: format {self.format!r} is not on the hot path")
        self.dtype = dtype
        self._coo = {}
        self._host = None       # dense ndarray or (pos, crd, vals)
        self._assignment = None
        self._kernel = None
        self._desc = None
        self._dev = None
        self._state = "new"
        self._plan = None            # slab plan of this matrix (learned schedules)
        self._schedule_choice = {}   # (op, K) -> (class, schedule) picked for it

    def _values_changed(self):
        """Drops what was derived from the values: the slab plan and the
        learned schedule choices."""
        self._plan = None
        self._schedule_choice = {}

    # -- values
    def insert(self, coords, value):
        if len(coords) != len(self.dims):
            raise TacoError(f"{self.name}: {len(coords)} coordinates for order {len(self.dims)}")
        # taco's insert appends to a coordinate buffer; pack sums the values
        # of repeated coordinates (test/tests-tensor.cpp:68-79)
        key = tuple(int(c) for c in coords)
        self._coo[key] = self._coo.get(key, 0.0) + value
        self._host = None
        self._values_changed()

    def pack(self):
        if self.format == CSR:
            m, n = self.dims
            keys = sorted(self._coo)
            counts = np.zeros(m, np.int64)
            for r, _ in keys:
                counts[r] += 1
            pos = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
            crd = np.array([c for _, c in keys], np.int32)
            vals = np.array([self._coo[k] for k in keys], self.dtype.np)
            self._host = (pos, crd, vals)
        else:
            d = np.zeros(self.dims, self.dtype.np)
            for k, v in self._coo.items():
                d[k] = v
            self._host = d
        self._dev = None
        self._values_changed()

    def set_csr(self, pos, crd, vals):
        """Host CSR arrays (taco's pack() result) for a CSR tensor."""
        if self.format != CSR:
            raise TacoError(f"{self.name} is not CSR")
        self._host = (np.ascontiguousarray(pos, np.int32), np.ascontiguousarray(crd, np.int32),
                      np.ascontiguousarray(vals, self.dtype.np))
        self._dev = None
        self._values_changed()

    def set_values(self, array):
        """Dense values for a dense tensor."""
        a = np.ascontiguousarray(array, self.dtype.np)
        if list(a.shape) != self.dims:
            raise TacoError(f"{self.name}: values of shape {a.shape}, dims {self.dims}")
        self._host = a
        self._dev = None
        self._values_changed()

    def _ensure_host(self):
        if self._host is None:
            self.pack()

    def _descriptor(self):
        import torch
        if self._dev is None:
            self._ensure_host()
            tdt = torch.float64 if self.dtype is Float64 else torch.float32
            if self.format == CSR:
                pos, crd, vals = self._host
                self._dev = (torch.from_numpy(pos).cuda(), torch.from_numpy(crd).cuda(),
                             torch.from_numpy(vals).to(tdt).cuda())
                self._desc = D.csr(*self._dev, tuple(self.dims), self.dtype.tag)
            else:
                self._dev = torch.from_numpy(self._host).to(tdt).cuda()
                self._desc = D.dense(self._dev, self.dtype.tag)
        return self._desc

    # -- index notation
    def __getitem__(self, ivars):
        ivars = ivars if isinstance(ivars, tuple) else (ivars,)
        return Access(self, ivars)

    def __setitem__(self, ivars, expr):
        ivars = ivars if isinstance(ivars, tuple) else (ivars,)
        self._assignment = Assignment(Access(self, ivars), expr)
        self._state = "defined"

    def get_assignment(self):
        if self._assignment is None:
            raise TacoError(f"{self.name} has no assignment")
        return self._assignment

    # -- compile / assemble / compute
    def compile(self, stmt=None):
        if not should_use_cuda_codegen():
            raise TacoError("GPU code generation is off (set_cuda_codegen_enabled); the "
                            "facade compiles through CodeGen_HIP only")
        stmt = stmt if stmt is not None else self.get_assignment().concretize()
        if stmt.assignment.lhs.tensor is not self:
            raise TacoError(f"statement {stmt!r} does not assign {self.name}")
        kind, a, b = _statement_kind(stmt.assignment)
        if _learned["enabled"] and not stmt.commands and kind in ("spmv", "spmm"):
            if a.tensor.dtype is not self.dtype or b.tensor.dtype is not self.dtype:
                raise TacoError("operands and result must share one datatype")
            self._kernel = _Learned(kind, a.tensor)
            self._operands = (a.tensor, b.tensor)
            self._kind = kind
            self._state = "compiled"
            return
        family, block, per, warp, rows, bound = _recognise(stmt, kind, a)
        op = {"spmv": _lib.OP_SPMV, "spmm": _lib.OP_SPMM, "spgemm": _lib.OP_SPGEMM}[kind]
        if a.tensor.dtype is not self.dtype or b.tensor.dtype is not self.dtype:
            raise TacoError("operands and result must share one datatype")
        spec = codegen.spec(op, family, self.dtype.csize, block, per, warp, rows, bound)
        try:
            self._kernel = codegen.Kernel(spec)
        except _lib.TacoHipError as e:
            raise TacoError(str(e)) from e
        self._operands = (a.tensor, b.tensor)
        self._kind = kind
        self._state = "compiled"

    def _output_descriptor(self):
        import torch
        tdt = torch.float64 if self.dtype is Float64 else torch.float32
        if self._kind == "spgemm":
            self._release_output()
            self._desc = D.csr_empty(tuple(self.dims), self.dtype.tag)
            self._dev = "library"
        else:
            self._dev = torch.zeros(self.dims, dtype=tdt, device="cuda")
            self._desc = D.dense(self._dev, self.dtype.tag)
        return self._desc

    def _release_output(self):
        if self._dev == "library" and self._desc is not None:
            L = _lib.lib()
            import ctypes
            for p in (self._desc.level_ptr(1, 0), self._desc.level_ptr(1, 1), self._desc.vals_ptr()):
                if p:
                    L.taco_hip_unified_free(ctypes.c_void_p(p))
            self._desc = None
            self._dev = None

    def assemble(self):
        if self._state not in ("compiled", "assembled", "computed"):
            raise TacoError(f"{self.name}: compile() must be called before assemble()")
        out = self._output_descriptor()
        a, b = self._operands
        if self._kind == "spgemm":
            self._kernel.assemble(out, a._descriptor(), b._descriptor())
        self._state = "assembled"

    def compute(self):
        import torch
        if self._state == "compiled":
            self.assemble()
        if self._state not in ("assembled", "computed"):
            raise TacoError(f"{self.name}: compile() must be called before compute()")
        a, b = self._operands
        self._kernel.compute(self._desc, a._descriptor(), b._descriptor())
        torch.cuda.synchronize()
        self._host = None
        self._values_changed()
        self._state = "computed"

    def evaluate(self):
        self.compile()
        self.assemble()
        self.compute()

    # -- results
    def to_numpy(self):
        """Dense values (dense tensors) as a host array."""
        if self.format == CSR:
            raise TacoError(f"{self.name} is CSR: use to_csr()")
        if self._state == "computed":
            return self._dev.cpu().numpy()
        self._ensure_host()
        return self._host

    def to_csr(self):
        """(pos, crd, vals) host arrays of a CSR tensor."""
        if self.format != CSR:
            raise TacoError(f"{self.name} is not CSR")
        if self._state == "computed" and self._dev == "library":
            import ctypes
            L = _lib.lib()
            nnz = self._desc.t.vals_size
            pos = np.empty(self.dims[0] + 1, np.int32)
            crd = np.empty(nnz, np.int32)
            vals = np.empty(nnz, self.dtype.np)
            for h, p in ((pos, self._desc.level_ptr(1, 0)), (crd, self._desc.level_ptr(1, 1)),
                         (vals, self._desc.vals_ptr())):
                if h.nbytes:
                    _lib.check(L.taco_hip_memcpy(h.ctypes.data_as(ctypes.c_void_p),
                                                 ctypes.c_void_p(p), h.nbytes, 1))
            return pos, crd, vals
        self._ensure_host()
        return self._host

    def __del__(self):
        try:
            self._release_output()
        except Exception:
            pass


# ------------------------------------------------------------- equals --

def _stored(t):
    """(coordinates as a (k, order) int64 array in storage order, values)."""
    if t.format == CSR:
        pos, crd, vals = t.to_csr()
        rows = np.repeat(np.arange(t.dims[0], dtype=np.int64), np.diff(pos))
        return np.stack([rows, crd.astype(np.int64)], axis=1), np.asarray(vals)
    a = np.asarray(t.to_numpy()).reshape(-1)
    coords = np.stack(np.unravel_index(np.arange(a.size), t.dims), axis=1).astype(np.int64)
    return coords.reshape(a.size, len(t.dims)), a


def _scalar_equals(a, b):
    """taco's scalarEquals: |(a - b) / a| must not exceed 1e-5, computed in
    double with a as the divisor (0/0 is NaN and passes; x/0 is inf and
    fails)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        diff = (np.asarray(a, np.float64) - np.asarray(b, np.float64)) / np.asarray(a, np.float64)
    return ~(np.abs(diff) > 10e-6)


def equals(a, b):
    """taco's tensor predicate (src/tensor.cpp:1038-1149): same component
    type, order and dimensions, then a merge over both tensors' stored
    components in storage order.  Equal coordinates need scalarEquals;
    a component stored in one tensor only must be zero.  Fill values are
    zero for every tensor here."""
    if a.dtype is not b.dtype or len(a.dims) != len(b.dims) or list(a.dims) != list(b.dims):
        return False
    ac, av = _stored(a)
    bc, bv = _stored(b)
    if ac.shape == bc.shape and np.array_equal(ac, bc):
        return bool(np.all(_scalar_equals(av, bv)))
    # general merge, in storage (lexicographic) order
    ia = ib = 0
    while ia < len(av) and ib < len(bv):
        ka, kb = tuple(ac[ia]), tuple(bc[ib])
        if ka != kb:
            if av[ia] == 0:
                ia += 1
                continue
            if bv[ib] == 0:
                ib += 1
                continue
            return False
        if not _scalar_equals(av[ia], bv[ib]):
            return False
        ia += 1
        ib += 1
    return bool(np.all(av[ia:] == 0) and np.all(bv[ib:] == 0))


# ----------------------------------------------------------------- files --

def read(filename, fmt=None, name=None):
    """taco::read (reference include/taco/tensor.h read(), file_io_*.cpp):
    a packed Tensor from a MatrixMarket (.mtx, .ttx) or FROSTT (.tns) file."""
    from . import io
    return io.read(filename, fmt, name)


def write(filename, tensor):
    """taco::write: MatrixMarket or FROSTT by the file's extension."""
    from . import io
    io.write(filename, tensor)
