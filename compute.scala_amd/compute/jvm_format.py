"""Text formatting of FP32 values the way the JVM prints them.

`Tensor.toString` joins `Float.toString` values (Tensors.scala:776-811). Float.toString
prints the shortest decimal that reads back to the same float, in plain notation for
1e-3 <= |x| < 1e7 and as d.dddE<exp> otherwise, always with a digit after the point.
"""
import numpy as np


def _shortest_digits(f32):
    """(sign, digits, exponent) with value = 0.d1d2... x 10^(exponent+1)."""
    s = np.format_float_scientific(np.float32(f32), unique=True, trim="-", exp_digits=1)
    sign = s.startswith("-")
    if sign:
        s = s[1:]
    mant, exp = s.split("e")
    digits = mant.replace(".", "")
    digits = digits.rstrip("0") or "0"
    return sign, digits, int(exp)


def float_to_string(value):
    f = np.float32(value)
    if np.isnan(f):
        return "NaN"
    if np.isinf(f):
        return "Infinity" if f > 0 else "-Infinity"
    if f == 0:
        return "-0.0" if np.signbit(f) else "0.0"
    sign, digits, exp = _shortest_digits(f)
    a = abs(float(f))
    prefix = "-" if sign else ""
    if 1e-3 <= a < 1e7:
        if exp >= 0:
            intpart = digits[:exp + 1].ljust(exp + 1, "0")
            frac = digits[exp + 1:] or "0"
        else:
            intpart = "0"
            frac = "0" * (-exp - 1) + digits
        return prefix + intpart + "." + frac
    frac = digits[1:] or "0"
    return prefix + digits[0] + "." + frac + "E" + str(exp)


def nested_to_string(flat, shape):
    """[a,b] nesting of Tensors.scala:783-808; a scalar prints bare."""
    flat = [float_to_string(v) for v in np.asarray(flat, dtype=np.float32).ravel()]
    shape = list(shape)
    if not shape:
        if len(flat) != 1:
            raise ValueError("shape() does not match the data size (%d)" % len(flat))
        return flat[0]

    def rec(dims, items):
        head = dims[0]
        if len(dims) == 1:
            if head != len(items):
                raise ValueError("shape does not match the data size")
            return "[" + ",".join(items) + "]"
        group = len(items) // head if head else 0
        return "[" + ",".join(rec(dims[1:], items[i * group:(i + 1) * group]) for i in range(head)) + "]"

    return rec(shape, flat)
