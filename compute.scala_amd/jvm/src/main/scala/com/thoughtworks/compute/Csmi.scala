package com.thoughtworks.compute

import java.nio.ByteBuffer

/** The native entry points of libcsmi.so, one per `cs_*` function the Tensor facade uses
  * (include/csmi.h), through the JNI shim `csmi_jni.c`. They replace the LWJGL OpenCL calls
  * beneath `trait OpenCL` (OpenCL.scala:1139-1433).
  *
  * Conventions (the shim's): handles and device pointers are `Long`; out-parameters are
  * one-element arrays; host memory is a direct `ByteBuffer`; every method returns the
  * `cs_status`, whose values are OpenCL's status codes, so [[checkErrorCode]] raises the
  * exception classes of `OpenCL.Exceptions` (OpenCL.scala:143-312).
  */
object Csmi {
  System.loadLibrary("csmi_jni")

  // status codes used by the facade (equal to CL_*)
  final val Success = 0
  final val InvalidValue = -30
  final val InvalidBufferSize = -61

  // cs_op (include/csmi.h)
  final val OpLiteral = 1
  final val OpParameter = 2
  final val OpTransform = 3
  final val OpFill = 4
  final val OpExtract = 5
  final val OpNeg = 10
  final val OpAbs = 11
  final val OpSqrt = 12
  final val OpTanh = 13
  final val OpExp = 14
  final val OpLog = 15
  final val OpAdd = 20
  final val OpSub = 21
  final val OpMul = 22
  final val OpDiv = 23
  final val OpMod = 24
  final val OpMin = 25
  final val OpMax = 26

  final val MaxRank = 8

  /** The cs_node record: field offsets in bytes and the record size. */
  object Node {
    final val Op = 0
    final val A = 4
    final val B = 8
    final val Param = 12
    final val Value = 16
    final val Rank = 20
    final val Shape = 24 // MaxRank int32 values
    final val Matrix = 56 // MaxRank * (MaxRank + 1) float64 values
    final val Size = 632
  }

  // math modes (cs_math_mode)
  final val MathPrecise = 0
  final val MathFast = 1
  final val MathAccurate = 2

  // GEMM variants (cs_gemm_variant)
  final val GemmFp32 = 0
  final val GemmBf16 = 1
  final val GemmBf16x3 = 2
  final val GemmFp8 = 3

  // cs_info_key
  final val InfoDeviceType = 1
  final val InfoComputeUnits = 2
  final val InfoMaxWorkGroupSize = 3

  @native def statusString(status: Int): String
  @native def lastError(): String

  @native def deviceCount(out: Array[Int]): Int
  @native def deviceOpen(ordinal: Int, out: Array[Long]): Int
  @native def deviceClose(device: Long): Int
  @native def deviceInfo(device: Long, key: Int, out: Array[Long]): Int
  @native def streamCreate(device: Long, out: Array[Long]): Int
  @native def streamDestroy(stream: Long): Int
  @native def streamSynchronize(stream: Long): Int

  @native def eventRecord(stream: Long, out: Array[Long]): Int
  @native def eventRelease(event: Long): Int
  @native def eventWait(event: Long): Int
  @native def eventElapsedMs(start: Long, end: Long, out: Array[Float]): Int
  @native def eventOnComplete(event: Long, continuation: Runnable): Int

  @native def bufferAllocOn(stream: Long, bytes: Long, out: Array[Long]): Int
  @native def bufferRetain(buffer: Long): Int
  @native def bufferRelease(buffer: Long): Int
  @native def bufferDevicePtr(buffer: Long, out: Array[Long]): Int
  @native def bufferWrite(stream: Long, buffer: Long, offset: Long, host: ByteBuffer, bytes: Long,
                          blocking: Boolean): Int
  @native def bufferRead(stream: Long, buffer: Long, offset: Long, host: ByteBuffer, bytes: Long): Int

  /** options: Array(mathMode, vectorWidth, itemsPerThread, blockSize, nontemporalStore[, launchOrder]);
    * launchOrder 0 (the default) tiles across a transposed leaf, 1 keeps the linear order. */
  @native def fusedCompile(device: Long, nodes: ByteBuffer, nnodes: Int, outputs: Array[Int], shape: Array[Int],
                           options: Array[Int], out: Array[Long]): Int
  @native def fusedCompileSum(device: Long, nodes: ByteBuffer, nnodes: Int, outputs: Array[Int],
                              shape: Array[Int], options: Array[Int], out: Array[Long]): Int
  @native def fusedLookup(device: Long, nodes: ByteBuffer, nnodes: Int, outputs: Array[Int], shape: Array[Int],
                          options: Array[Int], hit: Array[Int]): Int
  @native def fusedRelease(fused: Long): Int
  @native def fusedCacheClear(device: Long): Int
  @native def fusedLaunch(stream: Long, fused: Long, leaves: Array[Long], out: Long): Int
  @native def fusedSum(stream: Long, fused: Long, leaves: Array[Long], out: Long, out64: Long): Int

  @native def batchCreate(device: Long, exprs: Array[Long], block: Int, out: Array[Long]): Int
  @native def batchBind(stream: Long, batch: Long, args: Array[Long]): Int
  @native def batchLaunch(stream: Long, batch: Long): Int
  @native def batchRelease(batch: Long): Int
  @native def graphBeginCapture(stream: Long): Int
  @native def graphEndCapture(stream: Long, out: Array[Long]): Int
  @native def graphLaunch(stream: Long, graph: Long): Int
  @native def graphRelease(graph: Long): Int

  @native def reduceSum(stream: Long, in: Long, n: Long, out: Long, out64: Long): Int
  @native def reduceMax(stream: Long, in: Long, n: Long, out: Long): Int
  @native def reduceRows(stream: Long, in: Long, rows: Long, cols: Long, rowSum: Long, rowMax: Long, total: Long,
                         total64: Long): Int
  @native def reduceCols(stream: Long, in: Long, rows: Long, cols: Long, colSum: Long): Int
  @native def random(stream: Long, out: Long, n: Long, seed: Int): Int
  @native def randomNormal(stream: Long, out: Long, n: Long, seed: Int): Int
  @native def gemm(stream: Long, variant: Int, m: Int, n: Int, k: Int, a: Long, lda: Int, b: Long, ldb: Int,
                   c: Long, ldc: Int): Int

  /** Raises the exception class the reference raises for `status` (OpenCL.scala:143-312),
    * with the library's message for this thread. */
  def checkErrorCode(status: Int): Unit =
    if (status != Success) {
      OpenCL.checkErrorCode(status) // throws OpenCL.Exceptions.* for every CL status
      throw new IllegalStateException(statusString(status) + ": " + lastError())
    }
}
