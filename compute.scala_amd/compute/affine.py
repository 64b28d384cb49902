"""N-dimensional affine transforms on coordinates.

Mirror of NDimensionalAffineTransform (NDimensionalAffineTransform.scala:8-129). A matrix
mapping an n_in-dimensional space to an n_out-dimensional one is a flat list of
n_out * (n_in + 1) doubles, row-major; the last column is the translation.
"""


def translate(offset):
    """Translation matrix (NDimensionalAffineTransform.scala:13-26)."""
    n = len(offset)
    m = [0.0] * (n * (n + 1))
    for i in range(n):
        m[i * (n + 1) + i] = 1.0
        m[i * (n + 1) + n] = float(offset[i])
    return m


def _concatenate(m01, m12, length0, length1, length2):
    # (NDimensionalAffineTransform.scala:48-95)
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


def pre_concatenate(m01, m12, length0):
    """m12 after m01, with m01 : space0 -> space1 (NDimensionalAffineTransform.scala:28-35)."""
    length1 = len(m01) // (length0 + 1)
    length2 = len(m12) // (length1 + 1)
    return _concatenate(m01, m12, length0, length1, length2)


def concatenate(m12, m01, length2):
    """The same composition given the row count of m12 (NDimensionalAffineTransform.scala:37-46)."""
    length1 = len(m12) // length2 - 1
    length0 = len(m01) // length1 - 1
    return _concatenate(m01, m12, length0, length1, length2)


def transform(matrix, source):
    """Applies `matrix` to the point `source` (NDimensionalAffineTransform.scala:97-127)."""
    n = len(source)
    if len(matrix) % (n + 1) != 0:
        raise ValueError("matrix of %d entries does not take %d coordinates" % (len(matrix), n))
    rows = len(matrix) // (n + 1)
    out = []
    for y in range(rows):
        acc = matrix[y * (n + 1) + n]
        for x in range(n):
            acc = acc + matrix[y * (n + 1) + x] * source[x]
        out.append(acc)
    return out
