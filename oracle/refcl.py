"""refcl-host: the reference's generated OpenCL C kernels, run as host C (test infrastructure).

This is a restatement of OpenCLKernelBuilder's templates (OpenCLKernelBuilder.scala:135-571,
632-652) and of the export walk of Trees (Trees.scala:201-203, 220) and
parameterDescendants (Tensors.scala:230-251). For a closure DAG it emits the OpenCL C text
the reference would generate (`jit_kernel`, one `const` SSA statement per operation,
views as `(int)(...)` index expressions with bounds checks and padding), compiles that
text as host C with a small prelude and runs the NDRange over OpenMP threads. It stands
in for the POCL CPU device the reference targets; the arithmetic is the host compiler's
and libm's, in one of two modes:

  "precise": -ffp-contract=off (no FMA), libm float functions;
  "fast":    -funsafe-math-optimizations -ffp-contract=fast, the counterpart of the reference's
             -cl-unsafe-math-optimizations (OpenCL.scala:1131-1135).

Only tests, smoke checks and the CPU-baseline leg of the benchmark use this module.
The input DAG is the facade's node objects (duck-typed by class name); nothing from the
product's kernels or library is called.
"""
import ctypes
import decimal
import hashlib
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build", "refcl")
_lock = threading.Lock()


class BuildProgramFailure(Exception):
    """The host compiler rejected the kernel text (the reference's BuildProgramFailure)."""


# ----------------------------------------------------------------------------- literals
def java_float_to_string(value):
    """Float.toString: shortest round-trip digits; plain for 1e-3 <= |x| < 1e7."""
    f = np.float32(value)
    if np.isnan(f):
        return "NaN"
    if np.isinf(f):
        return "Infinity" if f > 0 else "-Infinity"
    if f == 0:
        return "-0.0" if np.signbit(f) else "0.0"
    s = np.format_float_scientific(f, unique=True, trim="-", exp_digits=1)
    neg = s.startswith("-")
    s = s.lstrip("-")
    mant, exp = s.split("e")
    exp = int(exp)
    digits = mant.replace(".", "").rstrip("0") or "0"
    a = abs(float(f))
    sign = "-" if neg else ""
    if 1e-3 <= a < 1e7:
        if exp >= 0:
            ip = digits[:exp + 1].ljust(exp + 1, "0")
            fp = digits[exp + 1:] or "0"
        else:
            ip, fp = "0", "0" * (-exp - 1) + digits
        return sign + ip + "." + fp
    return sign + digits[0] + "." + (digits[1:] or "0") + "E" + str(exp)


def float_literal(value):
    """ClFloatType.literal (OpenCLKernelBuilder.scala:251-264)."""
    f = np.float32(value)
    if np.isnan(f):
        return "NAN"
    if np.isinf(f):
        return "INFINITE" if f > 0 else "(-INFINITE)"
    return java_float_to_string(f) + "f"


def to_literal(value):
    """toLiteral: java.text.DecimalFormat with at most 3 fraction digits, HALF_EVEN on the
    exact binary value, grouping off (OpenCLKernelBuilder.scala:14-32)."""
    d = decimal.Decimal(float(value)).quantize(decimal.Decimal("0.001"), rounding=decimal.ROUND_HALF_EVEN)
    s = format(d, "f")
    if "." in s:
        s = s.rstrip("0").rstrip(".")
    if s in ("-0",):
        s = "-0"
    return s


# ----------------------------------------------------------------------------- emitter
def _concatenate(m12, m01, length2):
    """NDimensionalAffineTransform.concatenate (NDimensionalAffineTransform.scala:37-95)."""
    length1 = len(m12) // length2 - 1
    length0 = len(m01) // length1 - 1
    m02 = [0.0] * ((length0 + 1) * length2)
    for i2 in range(length2):
        for i0 in range(length0):
            acc = 0.0
            for i1 in range(length1):
                acc = acc + m12[i2 * (length1 + 1) + i1] * m01[i1 * (length0 + 1) + i0]
            m02[i2 * (length0 + 1) + i0] = acc
        acc = m12[i2 * (length1 + 1) + length1]
        for i1 in range(length1):
            acc = acc + m12[i2 * (length1 + 1) + i1] * m01[i1 * (length0 + 1) + length0]
        m02[i2 * (length0 + 1) + length0] = acc
    return m02


class _Emitter(object):
    def __init__(self):
        self.seed = 0
        self.global_definitions = []
        self.local = []
        self.type_cache = {}

    def fresh(self, prefix):
        """GlobalContext.freshName (OpenCLKernelBuilder.scala:86-98)."""
        enc = "".join(c if c.isalnum() and c.isascii() else "_" for c in prefix)
        name = ("_%s_%d" if (not enc or enc[0].isdigit()) else "%s_%d") % (enc, self.seed)
        self.seed += 1
        return name

    def array_type(self, shape):
        name = self.fresh("float_array")  # define() runs on every lookup
        key = ("array", tuple(shape))
        if key not in self.type_cache:
            self.type_cache[key] = name
            dims = "".join("[%d]" % s for s in shape)
            self.global_definitions.append("typedef global const float (* %s)%s;\n" % (name, dims))
        return self.type_cache[key]

    def tuple_type(self, length):
        name = self.fresh("float_tuple")
        key = ("tuple", length)
        if key not in self.type_cache:
            self.type_cache[key] = name
            self.global_definitions.append(
                "typedef struct %s { float tuple_data[%d]; } %s;\n" % (name, length, name))
        return self.type_cache[key]

    # terms: ("float", code) | ("param", code, type, shape, pad) | ("view", code, type, shape, pad, matrix)
    #        | ("fill", float_term)
    def export(self, node, memo):
        key = id(node)
        if key in memo:
            return memo[key]
        kind = type(node).__name__
        if kind == "FloatLiteral":
            t = ("float", float_literal(node.value))
        elif kind == "ArrayParameter":
            pad = float_literal(node.padding)
            typ = self.array_type(node.shape)
            code = self.fresh(str(memo["#ids"][key]))
            t = ("param", code, typ, tuple(node.shape), pad)
        elif kind == "Transform":
            arr = self.export(node.array, memo)
            if arr[0] == "param":
                t = ("view", arr[1], arr[2], arr[3], arr[4], list(node.matrix))
            elif arr[0] == "view":
                t = ("view", arr[1], arr[2], arr[3], arr[4],
                     _concatenate(arr[5], list(node.matrix), len(arr[3])))
            else:
                t = arr  # ArrayFill.transform returns itself
        elif kind == "Fill":
            t = ("fill", self.export(node.element, memo))
        elif kind == "Extract":
            t = self.extract(self.export(node.array, memo))
        elif kind == "Unary":
            a = self.export(node.operand, memo)[1]
            v = self.fresh("")
            expr = {"neg": "-%s", "abs": "fabs(%s)", "sqrt": "sqrt(%s)", "tanh": "tanh(%s)",
                    "exp": "exp(%s)", "log": "log(%s)"}[node.op] % a
            self.local.append("const float %s = %s;\n" % (v, expr))
            t = ("float", v)
        elif kind == "Binary":
            a = self.export(node.lhs, memo)[1]
            b = self.export(node.rhs, memo)[1]
            v = self.fresh("")
            if node.op in ("min", "max"):
                expr = "f%s(%s, %s)" % (node.op, a, b)
            else:
                expr = "%s %s %s" % (a, node.op, b)
            self.local.append("const float %s = %s;\n" % (v, expr))
            t = ("float", v)
        else:
            raise TypeError("unknown node " + kind)
        memo[key] = t
        return t

    def extract(self, arr):
        if arr[0] == "fill":
            return arr[1]
        if arr[0] == "param":
            # ArrayParameter.extract (OpenCLKernelBuilder.scala:435-455)
            _, code, _, shape, pad = arr
            idx = "".join("[__global_id_%d]" % i for i in range(len(shape)))
            bounds = ["__global_id_%d < %d" % (i, s) for i, s in enumerate(shape)]
            v = self.fresh("")
            deref = "(*%s)%s" % (code, idx)
            rhs = deref if not bounds else "(%s) ? %s : %s" % (" && ".join(bounds), deref, pad)
            self.local.append("const float %s = %s;\n" % (v, rhs))
            return ("float", v)
        # ArrayView.extract (OpenCLKernelBuilder.scala:348-411)
        _, code, _, shape, pad, matrix = arr
        rows = len(shape)
        if rows == 0:
            v = self.fresh("")
            self.local.append("const float %s = *%s;\n" % (v, code))
            return ("float", v)
        cols = len(matrix) // rows
        indices, bounds = [], []
        for y in range(rows):
            products = []
            for x in range(cols):
                scale = matrix[y * cols + x]
                if scale == 0.0:
                    continue
                if x < cols - 1:
                    products.append("__global_id_%d" % x if scale == 1.0 else
                                    "__global_id_%d * %s" % (x, to_literal(scale)))
                else:
                    products.append(to_literal(scale))
            ix = self.fresh("index")
            if not products:
                self.local.append("const int %s = 0;\n" % ix)
            else:
                self.local.append("const int %s = (int)(%s);\n" % (ix, " + ".join(products)))
                bounds += ["((bool)(%s >= 0))" % ix, "((bool)(%s < %d))" % (ix, shape[y])]
            indices.append(ix)
        v = self.fresh("")
        deref = "(*%s)%s" % (code, "".join("[%s]" % i for i in indices))
        rhs = deref if not bounds else "(%s) ? %s : %s" % (" && ".join(bounds), deref, pad)
        self.local.append("const float %s = %s;\n" % (v, rhs))
        return ("float", v)


def _parameters(closure):
    """parameterDescendants (Tensors.scala:230-251): first-visit DFS, by identity."""
    seen, out = set(), []

    def kids(n):
        k = type(n).__name__
        if k in ("Transform", "Extract"):
            return [n.array]
        if k == "Fill":
            return [n.element]
        if k == "Unary":
            return [n.operand]
        if k == "Binary":
            return [n.lhs, n.rhs]
        if k == "Concatenate":
            return list(n.elements)
        return []

    stack = [closure]
    while stack:
        n = stack.pop()
        if id(n) in seen:
            continue
        seen.add(id(n))
        if type(n).__name__ == "ArrayParameter":
            out.append(n)
            continue
        stack.extend(reversed(kids(n)))
    return out


def generate(closure, shape):
    """Returns (source, leaves, tuple_width) of the reference's kernel for `closure` over
    `shape` (enqueueClosure's loader, Tensors.scala:1297-1322)."""
    e = _Emitter()
    leaves = _parameters(closure)
    memo = {"#ids": {id(p): i for i, p in enumerate(leaves)}}
    is_tuple = type(closure).__name__ == "Concatenate"
    if is_tuple:
        elems = [e.export(x, memo)[1] for x in closure.elements]
        ttype = e.tuple_type(len(elems))
        body = e.fresh("")
        e.local.append("const %s %s = {{\n  %s\n}};\n" % (ttype, body, ",\n  ".join(elems)))
        out_type = ttype
    else:
        body = e.export(closure, memo)[1]
        out_type = "float"
    params = [memo[id(p)] for p in leaves]
    total = len(shape)
    cl_dims = 3
    out_id = e.fresh("output")
    decls = ["%s const restrict %s" % (p[2], p[1]) for p in params]
    out_param = "global %s * const restrict %s" % (out_type, out_id)
    if total > 0:
        idx = "__global_id_0"
        for d in range(1, total):
            idx = "(%s * __global_size_%d + __global_id_%d)" % (idx, d, d)
        assign = "%s[%s] = %s;\n" % (out_id, idx, body)
    else:
        assign = "*%s = %s;\n" % (out_id, body)
    if total <= cl_dims:
        dim_params = []
        dim_defs = "".join("const int __global_size_%d = get_global_size(%d);\n"
                           "const int __global_id_%d = get_global_id(%d);\n" % (d, d, d, d) for d in range(total))
    else:
        p = total - cl_dims
        dim_params = ["const int __global_size_%d" % d for d in range(1, p + 1)]
        strides = []
        stride = "1"
        for d in range(p, -1, -1):
            strides.append("const int __stride_%d = %s;\n" % (d, stride))
            stride = "__stride_%d * __global_size_%d" % (d, d)
        dim_defs = "".join(strides)
        dim_defs += "const int __global_size_0 = get_global_size(0) / __stride_0;\n"
        dim_defs += "const int __global_id_0 = get_global_id(0) / __stride_0;\n"
        dim_defs += "".join("const int __global_id_%d = get_global_id(0) / __stride_%d %% __global_size_%d;\n"
                            % (d, d, d) for d in range(1, p + 1))
        dim_defs += "".join("const int __global_size_%d = get_global_size(%d);\n"
                            "const int __global_id_%d = get_global_id(%d);\n" % (d + p, d, d + p, d)
                            for d in range(1, cl_dims))
    all_params = dim_params + decls + [out_param]
    src = "".join(e.global_definitions)
    src += "kernel void jit_kernel(%s) {\n%s%s%s}\n" % (", ".join(all_params), dim_defs, "".join(e.local), assign)
    width = len(closure.elements) if is_tuple else 1
    return src, leaves, params, out_type, width


_PRELUDE = r"""
#include <math.h>
#include <tgmath.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#define kernel static
#define global
typedef unsigned int uint;
static _Thread_local size_t refcl_gid[3];
static _Thread_local size_t refcl_gsize[3];
#define get_global_id(d) (refcl_gid[(d)])
#define get_global_size(d) (refcl_gsize[(d)])
"""


def _driver(params, out_type, total_dims):
    """refcl_run(gsize[3], bufs, dim_args): the NDRange over OpenMP threads."""
    p = max(total_dims - 3, 0)
    args = ["dim_args[%d]" % i for i in range(p)]
    args += ["(%s)bufs[%d]" % (prm[2], i) for i, prm in enumerate(params)]
    args.append("(%s *)bufs[%d]" % (out_type, len(params)))
    return r"""
void refcl_run(const int64_t *gs, void **bufs, const int *dim_args) {
  (void)dim_args;
  #pragma omp parallel for schedule(static)
  for (int64_t g0 = 0; g0 < gs[0]; ++g0) {
    refcl_gsize[0] = (size_t)gs[0]; refcl_gsize[1] = (size_t)gs[1]; refcl_gsize[2] = (size_t)gs[2];
    for (int64_t g1 = 0; g1 < gs[1]; ++g1)
      for (int64_t g2 = 0; g2 < gs[2]; ++g2) {
        refcl_gid[0] = (size_t)g0; refcl_gid[1] = (size_t)g1; refcl_gid[2] = (size_t)g2;
        jit_kernel(%s);
      }
  }
}
""" % ", ".join(args)


MODES = {
    "precise": ["-ffp-contract=off"],
    "fast": ["-funsafe-math-optimizations", "-ffp-contract=fast"],
}


class Kernel(object):
    """A compiled refcl-host kernel."""

    def __init__(self, closure, shape, mode="precise"):
        self.shape = tuple(int(s) for s in shape)
        self.source, self.leaves, params, out_type, self.width = generate(closure, self.shape)
        self.total = len(self.shape)
        code = _PRELUDE + self.source + _driver(params, out_type, self.total)
        digest = hashlib.sha1((code + mode + "|compile, then link without the mode flags").encode()).hexdigest()[:20]
        os.makedirs(BUILD, exist_ok=True)
        so = os.path.join(BUILD, digest + ".so")
        with _lock:
            if not os.path.exists(so):
                c = os.path.join(BUILD, digest + ".c")
                with open(c, "w") as f:
                    f.write(code)
                # compile with the mode's flags, link without them: gcc links crtfastmath.o into
                # a shared object built with -funsafe-math-optimizations, whose constructor sets
                # flush-to-zero and denormals-are-zero for the whole process on load (and the
                # FP64 oracle in the same process would then flush subnormals)
                obj = os.path.join(BUILD, digest + ".o")
                cmd = ["gcc", "-O2", "-std=gnu11", "-fopenmp", "-fPIC", "-c", "-w"] + MODES[mode] + ["-o", obj, c]
                proc = subprocess.run(cmd, capture_output=True, text=True)
                if proc.returncode != 0:
                    raise BuildProgramFailure(proc.stderr[-4000:])
                link = ["gcc", "-shared", "-fopenmp", "-o", so + ".tmp", obj, "-lm"]
                proc = subprocess.run(link, capture_output=True, text=True)
                if proc.returncode != 0:
                    raise BuildProgramFailure(proc.stderr[-4000:])
                os.replace(so + ".tmp", so)
        self._lib = ctypes.CDLL(so)
        self._lib.refcl_run.restype = None
        self._lib.refcl_run.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_void_p),
                                        ctypes.POINTER(ctypes.c_int)]

    def ndrange(self):
        """globalWorkSize of Tensors.scala:1346-1376 (a scalar launches as [1])."""
        s = self.shape
        if self.total == 0:
            return [1, 1, 1], []
        if self.total <= 3:
            g = list(s) + [1] * (3 - self.total)
            return g, []
        p = self.total - 3
        g0 = 1
        for d in range(p + 1):
            g0 *= s[d]
        return [g0, s[p + 1], s[p + 2]], [s[d] for d in range(1, p + 1)]

    def __call__(self, leaf_arrays):
        """Runs over `leaf_arrays` (float32, in leaf order); returns the flat output."""
        if len(leaf_arrays) != len(self.leaves):
            raise ValueError("kernel takes %d leaves" % len(self.leaves))
        n = 1
        for s in self.shape:
            n *= s
        out = np.zeros(n * self.width, dtype=np.float32)
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in leaf_arrays]
        bufs = (ctypes.c_void_p * (len(arrs) + 1))(*[a.ctypes.data for a in arrs], out.ctypes.data)
        g, dims = self.ndrange()
        if n == 0:
            return out
        gs = (ctypes.c_int64 * 3)(*g)
        da = (ctypes.c_int * max(len(dims), 1))(*dims)
        self._lib.refcl_run(gs, bufs, da)
        return out
