"""Expression DAG of tensor closures.

Mirror of Trees (Trees.scala:19-1054) restricted to what the Tensor API builds: float
literals and operators, array parameters (leaves), affine views (Transform), Fill,
Extract, and tuple Concatenate for `join`. Nodes are immutable; sharing a node object
shares the computation (the export memo of Trees.scala:220 works by identity).

`serialize` turns a closure into the cs_node array of include/csmi.h. Leaves are numbered
by first visit in depth-first, left-to-right order, which is both the alpha conversion
(Trees.scala:807-822) and the kernel argument order (Tensors.scala:230-251).
"""
import struct

from . import _capi as capi


class Tree(object):
    __slots__ = ()

    def children(self):
        return ()


class FloatLiteral(Tree):
    """float.literal (Trees.scala:373-383); compared by value bits."""
    __slots__ = ("value",)

    def __init__(self, value):
        self.value = struct.unpack("<f", struct.pack("<f", float(value)))[0]


class ArrayParameter(Tree):
    """array.parameter(id, padding, shape) (Trees.scala:754-823): a leaf buffer."""
    __slots__ = ("id", "padding", "shape")

    def __init__(self, ident, padding, shape):
        self.id = ident
        self.padding = struct.unpack("<f", struct.pack("<f", float(padding)))[0]
        self.shape = tuple(int(s) for s in shape)


class Transform(Tree):
    """array.transform(matrix) (Trees.scala:675-690): a view of `array`."""
    __slots__ = ("array", "matrix", "rank")

    def __init__(self, array, matrix, rank):
        self.array = array
        self.matrix = tuple(float(m) for m in matrix)
        self.rank = int(rank)  # rank of the view (columns - 1)

    def children(self):
        return (self.array,)


class Fill(Tree):
    """value.fill (Trees.scala:717-735): an array of one repeated float."""
    __slots__ = ("element",)

    def __init__(self, element):
        self.element = element

    def children(self):
        return (self.element,)


class Extract(Tree):
    """array.extract (Trees.scala:659-672): the element at the current coordinate."""
    __slots__ = ("array",)

    def __init__(self, array):
        self.array = array

    def children(self):
        return (self.array,)


UNARY = {"neg": capi.OP_NEG, "abs": capi.OP_ABS, "sqrt": capi.OP_SQRT, "tanh": capi.OP_TANH,
         "exp": capi.OP_EXP, "log": capi.OP_LOG}
BINARY = {"+": capi.OP_ADD, "-": capi.OP_SUB, "*": capi.OP_MUL, "/": capi.OP_DIV, "%": capi.OP_MOD,
          "min": capi.OP_MIN, "max": capi.OP_MAX}


class Unary(Tree):
    __slots__ = ("op", "operand")

    def __init__(self, op, operand):
        if op not in UNARY:
            raise ValueError("unknown unary op " + op)
        self.op = op
        self.operand = operand

    def children(self):
        return (self.operand,)


class Binary(Tree):
    __slots__ = ("op", "lhs", "rhs")

    def __init__(self, op, lhs, rhs):
        if op not in BINARY:
            raise ValueError("unknown binary op " + op)
        self.op = op
        self.lhs = lhs
        self.rhs = rhs

    def children(self):
        return (self.lhs, self.rhs)


class Concatenate(Tree):
    """tuple.join(elements...) (Trees.scala:953-973)."""
    __slots__ = ("elements",)

    def __init__(self, elements):
        self.elements = tuple(elements)

    def children(self):
        return self.elements


def parameters(tree):
    """Leaves in first-visit depth-first order, deduplicated by identity
    (parameterDescendants, Tensors.scala:230-251)."""
    seen = set()
    out = []
    stack = [tree]
    while stack:
        node = stack.pop()
        if id(node) in seen:
            continue
        seen.add(id(node))
        if isinstance(node, ArrayParameter):
            out.append(node)
            continue
        stack.extend(reversed(node.children()))
    return out


def serialize(tree):
    """Returns (nodes, outputs, leaves): a ctypes array of CsNode, the output node indices
    and the ArrayParameter leaves in slot order."""
    index = {}
    records = []
    leaves = []
    roots = tree.elements if isinstance(tree, Concatenate) else (tree,)

    def visit(root):
        # Iterative: pre-order numbers the leaves, post-order emits the nodes.
        stack = [(root, False)]
        while stack:
            node, done = stack.pop()
            key = id(node)
            if key in index:
                continue
            if not done:
                if isinstance(node, ArrayParameter):
                    leaves.append(node)
                    slot = len(leaves) - 1
                    index[key] = len(records)
                    records.append(("param", node, slot))
                    continue
                stack.append((node, True))
                for child in reversed(node.children()):
                    if id(child) not in index:
                        stack.append((child, False))
                continue
            index[key] = len(records)
            records.append(("node", node, None))

    for r in roots:
        if isinstance(r, Concatenate):
            raise ValueError("nested tuples are not supported")
        visit(r)

    nodes = (capi.CsNode * len(records))()
    for i, (kind, node, slot) in enumerate(records):
        n = nodes[i]
        n.a = n.b = -1
        if kind == "param":
            n.op = capi.OP_PARAMETER
            n.param = slot
            n.value = node.padding
            n.rank = len(node.shape)
            if n.rank > capi.MAX_RANK:
                raise ValueError("rank %d exceeds %d" % (n.rank, capi.MAX_RANK))
            for d, s in enumerate(node.shape):
                n.shape[d] = s
        elif isinstance(node, FloatLiteral):
            n.op = capi.OP_LITERAL
            n.value = node.value
        elif isinstance(node, Transform):
            n.op = capi.OP_TRANSFORM
            n.a = index[id(node.array)]
            n.rank = node.rank
            if node.rank > capi.MAX_RANK:
                raise ValueError("rank %d exceeds %d" % (node.rank, capi.MAX_RANK))
            for j, m in enumerate(node.matrix):
                n.matrix[j] = m
        elif isinstance(node, Fill):
            n.op = capi.OP_FILL
            n.a = index[id(node.element)]
        elif isinstance(node, Extract):
            n.op = capi.OP_EXTRACT
            n.a = index[id(node.array)]
        elif isinstance(node, Unary):
            n.op = UNARY[node.op]
            n.a = index[id(node.operand)]
        elif isinstance(node, Binary):
            n.op = BINARY[node.op]
            n.a = index[id(node.lhs)]
            n.b = index[id(node.rhs)]
        else:
            raise TypeError("cannot serialize %r" % (node,))
    outputs = [index[id(r)] for r in roots]
    return nodes, outputs, leaves


def structural_key(tree):
    """Canonical, injective structural identity of a closure: equal keys mean the same
    structure up to renaming of leaves (isSameStructure / structuralHashCode,
    Trees.scala:152-177, 278-298, 763-797, with the injective numbering the reference's
    one-way parameter map lacks)."""
    nodes, outputs, _ = serialize(tree)
    parts = []
    for n in nodes:
        row = (n.op, n.a, n.b, n.param if n.op == capi.OP_PARAMETER else -1,
               struct.pack("<f", n.value) if n.op in (capi.OP_LITERAL, capi.OP_PARAMETER) else b"",
               n.rank if n.op in (capi.OP_PARAMETER, capi.OP_TRANSFORM) else -1,
               tuple(n.shape[:n.rank]) if n.op == capi.OP_PARAMETER else (),
               tuple(n.matrix[:_matrix_len(nodes, n)]) if n.op == capi.OP_TRANSFORM else ())
        parts.append(row)
    return (tuple(parts), tuple(outputs), isinstance(tree, Concatenate))


def _matrix_len(nodes, n):
    src = nodes[n.a]
    rows = src.rank if src.op in (capi.OP_PARAMETER, capi.OP_TRANSFORM) else 0
    return rows * (n.rank + 1)


def is_same_structure(a, b):
    return structural_key(a) == structural_key(b)


def structural_hash(tree):
    return hash(structural_key(tree))
