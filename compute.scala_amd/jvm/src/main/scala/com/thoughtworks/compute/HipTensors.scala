package com.thoughtworks.compute

import java.nio.{ByteBuffer, ByteOrder}
import java.util.IdentityHashMap

import scala.collection.mutable

/** The Tensor API of `trait Tensors` (Tensors.scala:199-1442) on MI355X, over libcsmi through
  * [[Csmi]]. It takes the place of `trait Tensors extends OpenCL` for the `gpu` object
  * (gpu.scala:15-27), so that programs written against `import com.thoughtworks.compute.gpu._`
  * keep their source: the same tensor classes, operators, views, slow actions and companion
  * functions.
  *
  * What changes below the API: a closure is not exported as OpenCL C text. Its DAG is written
  * as `cs_node` records (first-visit order, leaves numbered as parameterDescendants does,
  * Tensors.scala:230-251) and handed to `cs_fused_compile`, whose structural cache replaces
  * `kernelCache` (Tensors.scala:1263-1289); leaves are bound by device pointer and the kernel
  * runs on this object's stream. Buffers come from the library's stream-ordered arena.
  *
  * This file mirrors the Python facade of this repository (compute/tensors.py), whose tests
  * pin every behaviour listed here; it is not compiled in the build image (no JDK there).
  */
trait HipTensors extends AutoCloseable {

  /** Device ordinal and math mode (Csmi.MathPrecise, MathFast or MathAccurate). */
  protected def deviceOrdinal: Int = 0
  protected def mathMode: Int = Csmi.MathPrecise

  private def check(status: Int): Unit = Csmi.checkErrorCode(status)

  private def one(f: Array[Long] => Int): Long = {
    val out = new Array[Long](1)
    check(f(out))
    out(0)
  }

  protected lazy val device: Long = one(Csmi.deviceOpen(deviceOrdinal, _))
  protected lazy val stream: Long = one(Csmi.streamCreate(device, _))

  private def fusedOptions: Array[Int] = Array(mathMode, 0, 0, 0, 0)

  def close(): Unit = {
    check(Csmi.streamDestroy(stream))
    check(Csmi.deviceClose(device))
  }

  // ------------------------------------------------------------------------------ buffers
  /** A counted reference to an arena block (PendingBuffer / DeviceBuffer,
    * Tensors.scala:253-299, OpenCL.scala:614-717). */
  final class DeviceBuffer private[HipTensors] (val handle: Long, val bytes: Long) extends AutoCloseable {
    val ptr: Long = one(Csmi.bufferDevicePtr(handle, _))
    def retain(): DeviceBuffer = {
      check(Csmi.bufferRetain(handle))
      new DeviceBuffer(handle, bytes)
    }
    def close(): Unit = check(Csmi.bufferRelease(handle))
  }

  private def allocate(bytes: Long): DeviceBuffer =
    new DeviceBuffer(one(Csmi.bufferAllocOn(stream, math.max(bytes, 1L), _)), bytes)

  private def directFloats(n: Int): ByteBuffer =
    ByteBuffer.allocateDirect(math.max(n, 1) * 4).order(ByteOrder.nativeOrder())

  // ------------------------------------------------------------------------------ trees
  /** The closure DAG (Trees.scala:303-1050): terms are compared by identity, as the export
    * walk memoises them (Trees.scala:220). */
  sealed trait Term
  final class FloatLiteral(val value: Float) extends Term
  final class ArrayParameter(val tensor: Tensor, val padding: Float, val shape: Array[Int]) extends Term
  final class Transform(val array: Term, val matrix: Array[Double], val rank: Int) extends Term
  final class Fill(val element: Term) extends Term
  final class Extract(val array: Term) extends Term
  final class Unary(val op: Int, val operand: Term) extends Term
  final class Binary(val op: Int, val lhs: Term, val rhs: Term) extends Term
  final class Concatenate(val elements: Seq[Term]) extends Term

  private def children(t: Term): Seq[Term] = t match {
    case x: Transform => Seq(x.array)
    case x: Fill => Seq(x.element)
    case x: Extract => Seq(x.array)
    case x: Unary => Seq(x.operand)
    case x: Binary => Seq(x.lhs, x.rhs)
    case x: Concatenate => x.elements
    case _ => Nil
  }

  /** cs_node records of `roots` (post-order, operands first), their output indices and the
    * leaves in slot order. Csmi.Node gives the record layout. */
  private final case class Serialized(nodes: ByteBuffer, count: Int, outputs: Array[Int], leaves: Seq[ArrayParameter])

  private def serialize(roots: Seq[Term]): Serialized = {
    val index = new IdentityHashMap[Term, Integer]
    val order = mutable.ArrayBuffer.empty[Term]
    val leaves = mutable.ArrayBuffer.empty[ArrayParameter]
    for (root <- roots) {
      val stack = mutable.Stack[(Term, Boolean)]((root, false))
      while (stack.nonEmpty) {
        val (node, done) = stack.pop()
        if (!index.containsKey(node)) {
          node match {
            case p: ArrayParameter =>
              leaves += p
              index.put(node, order.length)
              order += node
            case _ if !done =>
              stack.push((node, true))
              for (c <- children(node).reverse if !index.containsKey(c)) stack.push((c, false))
            case _ =>
              index.put(node, order.length)
              order += node
          }
        }
      }
    }
    val buf = ByteBuffer.allocateDirect(order.length * Csmi.Node.Size).order(ByteOrder.LITTLE_ENDIAN)
    def idx(t: Term): Int = index.get(t).intValue
    for ((node, i) <- order.zipWithIndex) {
      val base = i * Csmi.Node.Size
      def put(op: Int, a: Int, b: Int): Unit = {
        buf.putInt(base + Csmi.Node.Op, op)
        buf.putInt(base + Csmi.Node.A, a)
        buf.putInt(base + Csmi.Node.B, b)
      }
      node match {
        case p: ArrayParameter =>
          if (p.shape.length > Csmi.MaxRank) throw new IllegalArgumentException("rank exceeds " + Csmi.MaxRank)
          put(Csmi.OpParameter, -1, -1)
          buf.putInt(base + Csmi.Node.Param, leaves.indexWhere(_ eq p))
          buf.putFloat(base + Csmi.Node.Value, p.padding)
          buf.putInt(base + Csmi.Node.Rank, p.shape.length)
          for ((s, d) <- p.shape.zipWithIndex) buf.putInt(base + Csmi.Node.Shape + 4 * d, s)
        case l: FloatLiteral =>
          put(Csmi.OpLiteral, -1, -1)
          buf.putFloat(base + Csmi.Node.Value, l.value)
        case t: Transform =>
          put(Csmi.OpTransform, idx(t.array), -1)
          buf.putInt(base + Csmi.Node.Rank, t.rank)
          for ((m, j) <- t.matrix.zipWithIndex) buf.putDouble(base + Csmi.Node.Matrix + 8 * j, m)
        case f: Fill => put(Csmi.OpFill, idx(f.element), -1)
        case e: Extract => put(Csmi.OpExtract, idx(e.array), -1)
        case u: Unary => put(u.op, idx(u.operand), -1)
        case b: Binary => put(b.op, idx(b.lhs), idx(b.rhs))
        case _: Concatenate => throw new IllegalArgumentException("nested tuples are not supported")
      }
    }
    Serialized(buf, order.length, roots.map(idx).toArray, leaves.toList)
  }

  // ------------------------------------------------------------------------------ views
  /** The view of slice `index` along dimension d of a rank-n tensor (Tensors.scala:1038-1073). */
  private def splitMatrix(n: Int, d: Int, index: Int): Array[Double] = {
    val m = new Array[Double](n * n)
    for (i <- 0 until d) m(i * n + i) = 1.0
    m(d * n + n - 1) = index.toDouble
    for (i <- d + 1 until n) m(i * n + i - 1) = 1.0
    m
  }

  /** The view m01 (space0 -> space1, length0 = rank columns plus the constant) followed by the
    * checkpoint's m12 (space1 -> space2): NDimensionalAffineTransform's preConcatenate, in
    * its order of additions (NDimensionalAffineTransform.scala:28-95), which fixes the bits
    * of fractional coefficients. */
  private def preConcatenate(m01: Array[Double], m12: Array[Double], length0: Int): Array[Double] = {
    val length1 = m01.length / (length0 + 1)
    val length2 = m12.length / (length1 + 1)
    val m02 = new Array[Double]((length0 + 1) * length2)
    for (i2 <- 0 until length2) {
      for (i0 <- 0 until length0) {
        var acc = 0.0
        for (i1 <- 0 until length1) acc = acc + m12(i2 * (length1 + 1) + i1) * m01(i1 * (length0 + 1) + i0)
        m02(i2 * (length0 + 1) + i0) = acc
      }
      var acc = m12(i2 * (length1 + 1) + length1)
      for (i1 <- 0 until length1) acc = acc + m12(i2 * (length1 + 1) + i1) * m01(i1 * (length0 + 1) + length0)
      m02(i2 * (length0 + 1) + length0) = acc
    }
    m02
  }

  // ------------------------------------------------------------------------------ tensors
  /** Base of every tensor (Tensors.scala:636-1261). */
  sealed trait Tensor {
    def shape: Array[Int]
    def padding: Float
    def closure: Term

    private[HipTensors] def materialize(memo: IdentityHashMap[Tensor, DeviceBuffer]): DeviceBuffer

    private[HipTensors] def buffer(memo: IdentityHashMap[Tensor, DeviceBuffer]): DeviceBuffer = {
      var b = memo.get(this)
      if (b == null) {
        b = materialize(memo)
        memo.put(this, b)
      }
      b.retain()
    }

    /** array.parameter(this, padding, shape), once per tensor (Tensors.scala:1253-1260). */
    lazy val arrayTerm: ArrayParameter = new ArrayParameter(this, padding, shape)

    def doBuffer(): DeviceBuffer = {
      val memo = new IdentityHashMap[Tensor, DeviceBuffer]
      try buffer(memo)
      finally memo.values.forEach(_.close())
    }

    def flatArray: Array[Float] = {
      val n = shape.product
      val b = doBuffer()
      try {
        val host = directFloats(n)
        if (n > 0) check(Csmi.bufferRead(stream, b.handle, 0, host, n.toLong * 4))
        val out = new Array[Float](n)
        host.asFloatBuffer().get(out)
        out
      } finally b.close()
    }

    override def toString: String = JvmFormat.nested(flatArray, shape)

    /** Eagerly materialises into a CachedTensor that owns its buffer until `close()`
      * (doCache, Tensors.scala:642-666, without the RAII `Do` wrapper). */
    def doCache(): CachedTensor = new CachedTensor(shape, padding, doBuffer())

    def nonInline: Tensor = this

    protected def derived(c: Term): Tensor = new InlineTensor(shape, padding, c)

    private def binary(op: Int, rhs: Tensor): Tensor = {
      val (l, r) = autoBroadcast(this, rhs)
      new InlineTensor(l.shape, l.padding, new Binary(op, l.closure, r.closure))
    }

    def +(rhs: Tensor): Tensor = binary(Csmi.OpAdd, rhs)
    def -(rhs: Tensor): Tensor = binary(Csmi.OpSub, rhs)
    def *(rhs: Tensor): Tensor = binary(Csmi.OpMul, rhs)
    def /(rhs: Tensor): Tensor = binary(Csmi.OpDiv, rhs)
    def %(rhs: Tensor): Tensor = binary(Csmi.OpMod, rhs)
    def unary_- : Tensor = derived(new Unary(Csmi.OpNeg, closure))
    def unary_+ : Tensor = this

    /** Old dimension i maps to new dimension i; new trailing dimensions are appended and
      * size-1 dimensions stretch (Tensors.scala:816-855). */
    def broadcast(newShape: Array[Int]): Tensor = {
      if (newShape.sameElements(shape)) return this
      this match {
        case f: FillTensor => return new FillTensor(newShape, padding, f.value)
        case _ =>
      }
      val n = shape.length
      if (n > newShape.length)
        throw new IllegalArgumentException(s"Cannot broadcast ${shape.mkString("[", ",", "]")}")
      val m = new Array[Double]((newShape.length + 1) * n)
      for (i <- 0 until n) {
        if (shape(i) == newShape(i)) m(i * (newShape.length + 1) + i) = 1.0
        else if (shape(i) != 1)
          throw new IllegalArgumentException(
            s"Cannot broadcast ${shape.mkString("[", ",", "]")} to ${newShape.mkString("[", ",", "]")}")
      }
      transform(newShape, m)
    }

    /** A fusion barrier sharing this tensor's buffer (Tensors.scala:879-888). */
    def reshape(newShape: Array[Int]): Tensor = {
      if (newShape.product != shape.product)
        throw new IllegalArgumentException(s"reshape from ${shape.mkString(",")} to ${newShape.mkString(",")}")
      new ForwardingTensor(newShape, padding, this)
    }

    def scale(newShape: Array[Int]): Tensor = {
      val n = newShape.length
      if (n != shape.length) throw new IllegalArgumentException("scale must keep the rank")
      val m = new Array[Double](n * (n + 1))
      for (i <- 0 until n) m(i * (n + 1) + i) = shape(i).toDouble / newShape(i)
      transform(newShape, m)
    }

    /** out[i] = this[i - offset], padding outside (Tensors.scala:970-976). */
    def translate(offset: Array[Double], newShape: Array[Int] = shape): Tensor = {
      val n = offset.length
      if (n != shape.length) throw new IllegalArgumentException("offset rank differs from the tensor rank")
      val m = new Array[Double](n * (n + 1))
      for (i <- 0 until n) {
        m(i * (n + 1) + i) = 1.0
        m(i * (n + 1) + n) = -offset(i)
      }
      transform(newShape, m)
    }

    /** Views of a view compose into one matrix per checkpoint (Tensors.scala:978-1003). */
    def transform(newShape: Array[Int], matrix: Array[Double]): Tensor = this match {
      case t: TransformedTensor =>
        new TransformedTensor(newShape, padding, t.checkpoint, preConcatenate(matrix, t.matrix, newShape.length))
      case _ => new TransformedTensor(newShape, padding, this, matrix)
    }

    /** New dimension i is old dimension dimensions(i) (Tensors.scala:1008-1025). */
    def permute(dimensions: Array[Int]): Tensor = {
      val n = shape.length
      if (dimensions.length != n) throw new IllegalArgumentException(s"permute needs $n dimensions")
      val m = new Array[Double](n * (n + 1))
      for (i <- 0 until n) m(dimensions(i) * (n + 1) + i) = 1.0
      transform(dimensions.map(shape(_)), m)
    }

    def transpose: Tensor = permute(shape.indices.reverse.toArray)

    /** Views along `dimension`, without copies (Tensors.scala:1035-1074). */
    def split(dimension: Int): IndexedSeq[Tensor] = {
      if (dimension < 0 || dimension >= shape.length) throw new IllegalArgumentException("split dimension")
      val newShape = shape.patch(dimension, Nil, 1)
      (0 until shape(dimension)).map(i => transform(newShape, splitMatrix(shape.length, dimension, i)))
    }

    /** Full FP32 sum (Tensors.scala:673-771): one kernel that computes and adds the elements
      * of an inline tensor, or the reduction of a materialised one. */
    def sum: Tensor = new ReduceTensor(this)
  }

  /** (Tensors.scala:1400-1411) */
  class InlineTensor(val shape: Array[Int], val padding: Float, c: Term) extends Tensor {
    def closure: Term = c
    private[HipTensors] def materialize(memo: IdentityHashMap[Tensor, DeviceBuffer]): DeviceBuffer =
      enqueueClosure(Seq(closure), shape, memo)
    override def nonInline: Tensor = new ForwardingTensor(shape, padding, this)
  }

  /** (Tensors.scala:1413-1428) closure = checkpoint.arrayTerm.transform(matrix).extract */
  final class TransformedTensor(shape: Array[Int], padding: Float, val checkpoint: Tensor, val matrix: Array[Double])
      extends InlineTensor(shape, padding, null) {
    override lazy val closure: Term = new Extract(new Transform(checkpoint.arrayTerm, matrix, shape.length))
  }

  /** (Tensors.scala:1394-1397) */
  final class FillTensor(shape: Array[Int], padding: Float, val value: Float)
      extends InlineTensor(shape, padding, new FloatLiteral(value))

  /** (Tensors.scala:1431-1440) closure = arrayTerm.extract */
  sealed abstract class NonInlineTensor extends Tensor {
    lazy val closure: Term = new Extract(arrayTerm)
  }

  /** A tensor whose buffer is kept until `close()` (Tensors.scala:602, 642-666). */
  final class CachedTensor(val shape: Array[Int], val padding: Float, private var held: DeviceBuffer)
      extends NonInlineTensor
      with AutoCloseable {
    private[HipTensors] def materialize(memo: IdentityHashMap[Tensor, DeviceBuffer]): DeviceBuffer = {
      if (held == null) throw new IllegalStateException("the cache of this tensor has been released")
      held.retain()
    }
    def close(): Unit = if (held != null) { held.close(); held = null }
  }

  /** nonInline and reshape: a barrier sharing the source buffer (Tensors.scala:879-888, 1405-1410). */
  final class ForwardingTensor(val shape: Array[Int], val padding: Float, source: Tensor) extends NonInlineTensor {
    private[HipTensors] def materialize(memo: IdentityHashMap[Tensor, DeviceBuffer]): DeviceBuffer =
      source.buffer(memo)
  }

  /** Tensor(elements): uploaded through pinned staging (Tensors.scala:445-463). */
  final class DataTensor(val shape: Array[Int], val padding: Float, data: Array[Float]) extends NonInlineTensor {
    private[HipTensors] def materialize(memo: IdentityHashMap[Tensor, DeviceBuffer]): DeviceBuffer = {
      val b = allocate(data.length.toLong * 4)
      if (data.nonEmpty) {
        val host = directFloats(data.length)
        host.asFloatBuffer().put(data)
        check(Csmi.bufferWrite(stream, b.handle, 0, host, data.length.toLong * 4, true))
      }
      b
    }
  }

  /** Tensor.random / Tensor.randomNormal (Tensors.scala:479-524): the Wang hash on the device. */
  final class RandomTensor(val shape: Array[Int], val padding: Float, seed: Int, normal: Boolean)
      extends NonInlineTensor {
    private[HipTensors] def materialize(memo: IdentityHashMap[Tensor, DeviceBuffer]): DeviceBuffer = {
      val n = shape.product.toLong
      if (normal) {
        val b = allocate((n + (n & 1)) * 4) // pairs, as Tensors.scala:509-513 pads
        check(Csmi.randomNormal(stream, b.ptr, n, seed))
        b
      } else {
        val b = allocate(n * 4)
        check(Csmi.random(stream, b.ptr, n, seed))
        b
      }
    }
  }

  /** Tensor.join(seq): one kernel writing n floats per coordinate (Tensors.scala:577-598). */
  final class JoinTensor(parts: Seq[Tensor]) extends NonInlineTensor {
    val shape: Array[Int] = parts.head.shape :+ parts.length
    val padding: Float = parts.head.padding
    private[HipTensors] def materialize(memo: IdentityHashMap[Tensor, DeviceBuffer]): DeviceBuffer =
      enqueueClosure(parts.map(_.closure), parts.head.shape, memo)
  }

  /** `sum`: fused into an inline producer (one kernel, no intermediate buffer), otherwise the
    * reduction of the materialised source; FP64 combine in a fixed order. */
  final class ReduceTensor(source: Tensor) extends NonInlineTensor {
    val shape: Array[Int] = Array.empty
    val padding: Float = source.padding
    private[HipTensors] def materialize(memo: IdentityHashMap[Tensor, DeviceBuffer]): DeviceBuffer = {
      val out = allocate(4)
      source match {
        case s: InlineTensor if !memo.containsKey(s) && !s.closure.isInstanceOf[Concatenate] =>
          val ser = serialize(Seq(s.closure))
          val f = one(Csmi.fusedCompileSum(device, ser.nodes, ser.count, ser.outputs, s.shape, fusedOptions, _))
          try {
            val args = ser.leaves.map(_.tensor.buffer(memo))
            try check(Csmi.fusedSum(stream, f, args.map(_.ptr).toArray, out.ptr, 0L))
            finally args.foreach(_.close())
          } finally check(Csmi.fusedRelease(f))
        case _ =>
          val src = source.buffer(memo)
          try check(Csmi.reduceSum(stream, src.ptr, source.shape.product.toLong, out.ptr, 0L))
          finally src.close()
      }
      out
    }
  }

  /** C = A . B on MFMA (the reference writes it as split/broadcast/join idioms,
    * README.md:301-362). */
  final class GemmTensor(a: Tensor, b: Tensor, variant: Int) extends NonInlineTensor {
    if (a.shape.length != 2 || b.shape.length != 2 || a.shape(1) != b.shape(0))
      throw new IllegalArgumentException("matrixMultiply of incompatible shapes")
    val shape: Array[Int] = Array(a.shape(0), b.shape(1))
    val padding: Float = a.padding
    private[HipTensors] def materialize(memo: IdentityHashMap[Tensor, DeviceBuffer]): DeviceBuffer = {
      val Array(m, k) = a.shape
      val n = b.shape(1)
      val ba = a.buffer(memo)
      val bb = b.buffer(memo)
      try {
        val out = allocate(math.max(m.toLong * n, 1L) * 4)
        check(Csmi.gemm(stream, variant, m, n, k, ba.ptr, k, bb.ptr, n, out.ptr, n))
        out
      } finally {
        ba.close()
        bb.close()
      }
    }
  }

  /** autoBroadcast / autoBroadcastShape (Tensors.scala:202-222): leading dimensions align. */
  private def autoBroadcast(a: Tensor, b: Tensor): (Tensor, Tensor) = {
    val s1 = a.shape
    val s2 = b.shape
    val shape = (0 until math.max(s1.length, s2.length)).map { i =>
      if (i >= s1.length || s1(i) == 1) s2(i)
      else if (i >= s2.length || s2(i) == 1) s1(i)
      else if (s1(i) == s2(i)) s1(i)
      else throw new IllegalArgumentException(s"Cannot broadcast ${s1.mkString(",")} and ${s2.mkString(",")}")
    }.toArray
    (a.broadcast(shape), b.broadcast(shape))
  }

  /** enqueueClosure (Tensors.scala:1291-1392): compile or fetch from the library's structural
    * cache, materialise the leaves, allocate the output, launch on this object's stream. */
  private def enqueueClosure(roots: Seq[Term], shape: Array[Int],
                             memo: IdentityHashMap[Tensor, DeviceBuffer]): DeviceBuffer = {
    val ser = serialize(roots)
    val f = one(Csmi.fusedCompile(device, ser.nodes, ser.count, ser.outputs, shape, fusedOptions, _))
    try {
      val args = ser.leaves.map(_.tensor.buffer(memo))
      try {
        val out = allocate(math.max(shape.product.toLong, 1L) * roots.length * 4)
        check(Csmi.fusedLaunch(stream, f, args.map(_.ptr).toArray, out.ptr))
        out
      } finally args.foreach(_.close())
    } finally check(Csmi.fusedRelease(f))
  }

  // ------------------------------------------------------------------------------ companion
  /** `object Tensor` of the reference (Tensors.scala:395-600). */
  object Tensor {
    def apply(elements: Array[Float]): Tensor = new DataTensor(Array(elements.length), 0.0f, elements.clone())
    def apply(elements: Array[Array[Float]]): Tensor = {
      val cols = if (elements.isEmpty) 0 else elements.head.length
      if (elements.exists(_.length != cols)) throw new IllegalArgumentException("ragged nested sequence")
      new DataTensor(Array(elements.length, cols), 0.0f, elements.flatten)
    }
    def apply(value: Float): Tensor = new DataTensor(Array.empty, 0.0f, Array(value))
    def fill(value: Float, shape: Array[Int], padding: Float = 0.0f): Tensor = new FillTensor(shape, padding, value)
    def scalar(value: Float, padding: Float = 0.0f): Tensor = fill(value, Array.empty, padding)
    def random(shape: Array[Int], seed: Int = scala.util.Random.nextInt(), padding: Float = 0.0f): Tensor =
      new RandomTensor(shape, padding, seed, normal = false)
    def randomNormal(shape: Array[Int], seed: Int = scala.util.Random.nextInt(), padding: Float = 0.0f): Tensor =
      new RandomTensor(shape, padding, seed, normal = true)
    def join(tensors: Seq[Tensor]): Tensor = {
      if (tensors.isEmpty) throw new IllegalArgumentException("join of nothing")
      new JoinTensor(tensors)
    }
    def join(tensors: Seq[Tensor], dimension: Int): Tensor = {
      val joined = join(tensors)
      val rank = joined.shape.length
      if (dimension == rank - 1) joined
      else joined.permute((0 until rank).map(i => if (i < dimension) i else if (i == dimension) rank - 1 else i - 1).toArray)
    }
    def matrixMultiply(a: Tensor, b: Tensor, variant: Int = Csmi.GemmFp32): Tensor = new GemmTensor(a, b, variant)

    private def unary(op: Int, x: Tensor): Tensor = new InlineTensor(x.shape, x.padding, new Unary(op, x.closure))
    def abs(x: Tensor): Tensor = unary(Csmi.OpAbs, x)
    def sqrt(x: Tensor): Tensor = unary(Csmi.OpSqrt, x)
    def tanh(x: Tensor): Tensor = unary(Csmi.OpTanh, x)
    def exp(x: Tensor): Tensor = unary(Csmi.OpExp, x)
    def log(x: Tensor): Tensor = unary(Csmi.OpLog, x)
    def min(a: Tensor, b: Tensor): Tensor = {
      val (l, r) = autoBroadcast(a, b)
      new InlineTensor(l.shape, l.padding, new Binary(Csmi.OpMin, l.closure, r.closure))
    }
    def max(a: Tensor, b: Tensor): Tensor = {
      val (l, r) = autoBroadcast(a, b)
      new InlineTensor(l.shape, l.padding, new Binary(Csmi.OpMax, l.closure, r.closure))
    }
  }
}

/** JVM `Float.toString` of nested arrays, the reference's `toString` (Tensors.scala:776-811). */
object JvmFormat {
  def nested(flat: Array[Float], shape: Array[Int]): String = {
    def go(offset: Int, dim: Int): String =
      if (dim == shape.length) flat(offset).toString
      else {
        val stride = shape.drop(dim + 1).product
        (0 until shape(dim)).map(i => go(offset + i * stride, dim + 1)).mkString("[", ",", "]")
      }
    go(0, 0)
  }
}

/** The MI355X `gpu` object: `import com.thoughtworks.compute.gpu._` as before (gpu.scala:15-27),
  * GPU 0, precise math (the mode whose results TensorsSpec's known answers fix). */
object gpu extends HipTensors
