"""The Scala facade and JNI shim sources (compute.scala_amd/jvm/) against the C ABI, without
a JDK (CPU): the cs_node record layout the Scala serializer writes equals the library's, every
@native method has its JNI function and the reverse, every cs_* call of the shim is declared in
include/csmi.h and exported by libcsmi.so, and the constants equal the header's."""
import ctypes
import os
import re

from compute import _capi as capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JVM = os.path.join(ROOT, "compute.scala_amd", "jvm")
SCALA = os.path.join(JVM, "src", "main", "scala", "com", "thoughtworks", "compute")


def _read(*parts):
    with open(os.path.join(*parts)) as f:
        return f.read()


def test_node_layout_matches_the_library():
    src = _read(SCALA, "Csmi.scala")
    block = src[src.index("object Node {"):]
    block = block[:block.index("}")]
    got = {k: int(v) for k, v in re.findall(r"final val (\w+) = (\d+)", block)}
    fields = {f[0]: getattr(capi.CsNode, f[0]).offset for f in capi.CsNode._fields_}
    want = {"Op": fields["op"], "A": fields["a"], "B": fields["b"], "Param": fields["param"],
            "Value": fields["value"], "Rank": fields["rank"], "Shape": fields["shape"],
            "Matrix": fields["matrix"], "Size": ctypes.sizeof(capi.CsNode)}
    assert got == want
    assert want["Size"] == 632


def test_constants_match_the_header():
    src = _read(SCALA, "Csmi.scala")
    vals = {k: int(v) for k, v in re.findall(r"final val (\w+) = (-?\d+)", src)}
    ops = {"OpLiteral": 1, "OpParameter": capi.OP_PARAMETER, "OpTransform": capi.OP_TRANSFORM,
           "OpFill": capi.OP_FILL, "OpExtract": capi.OP_EXTRACT, "OpNeg": capi.OP_NEG, "OpAbs": capi.OP_ABS,
           "OpSqrt": capi.OP_SQRT, "OpTanh": capi.OP_TANH, "OpExp": capi.OP_EXP, "OpLog": capi.OP_LOG,
           "OpAdd": capi.OP_ADD, "OpSub": capi.OP_SUB, "OpMul": capi.OP_MUL, "OpDiv": capi.OP_DIV,
           "OpMod": capi.OP_MOD, "OpMin": capi.OP_MIN, "OpMax": capi.OP_MAX,
           "MaxRank": capi.MAX_RANK, "MathPrecise": capi.MATH_PRECISE, "MathFast": capi.MATH_FAST,
           "MathAccurate": capi.MATH_ACCURATE, "GemmFp32": capi.GEMM_FP32, "GemmBf16": capi.GEMM_BF16,
           "GemmBf16x3": capi.GEMM_BF16X3, "GemmFp8": capi.GEMM_FP8, "InvalidValue": -30,
           "InvalidBufferSize": -61, "Success": 0}
    for k, v in ops.items():
        assert vals[k] == v, k
    header = _read(ROOT, "include", "csmi.h")
    assert re.search(r"CS_OP_LITERAL = 1\b", header)


def test_natives_and_jni_functions_correspond():
    scala = _read(SCALA, "Csmi.scala")
    natives = set(re.findall(r"@native def (\w+)\(", scala))
    shim = _read(JVM, "csmi_jni.c")
    jni = set(re.findall(r"JNIEXPORT \w+ J\((\w+)\)", shim))
    assert natives == jni, (natives ^ jni)
    # every method the facade trait calls exists
    facade = _read(SCALA, "HipTensors.scala")
    called = set(re.findall(r"Csmi\.(\w+)\(", facade))
    assert called <= natives | {"checkErrorCode"}, called - natives


def test_shim_calls_exported_entry_points():
    shim = _read(JVM, "csmi_jni.c")
    header = _read(ROOT, "include", "csmi.h")
    calls = set(re.findall(r"\b(cs_[a-z0-9_]+)\(", shim)) - {"cs_status"}
    declared = set(re.findall(r"\b(cs_[a-z0-9_]+)\(", header))
    assert calls <= declared, calls - declared
    for name in calls:
        assert hasattr(capi.lib, name), name
