// Fused elementwise + view kernels: canonical DAG key, HIP code generation, the structural
// kernel cache and launch.
//
// This is the MI355X code generator that takes the place of OpenCLKernelBuilder
// (OpenCLKernelBuilder.scala:135-571) and of enqueueClosure's compile/cache/launch steps
// (Tensors.scala:1263-1392). It keeps the reference's method: one kernel per expression
// DAG, one SSA statement per operation, every view folded into the index arithmetic of
// its leaf, padding outside the source bounds. What changes is the shape of the kernel:
// a flattened 1-D grid where each thread produces `V` consecutive innermost outputs
// (`U` times), contiguous leaves are read with one wide load, broadcast leaves with one
// scalar load, and the outputs are written with wide coalesced stores.
#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <sstream>
#include <unordered_map>

#include <hip/hip_ext.h>

#include "internal.h"
#include "precise_math.inc"  // kPreciseMathSource, generated from precise_math.h


namespace csmi {
namespace {

uint64_t fnv1a(const std::string &s) {
  uint64_t h = 1469598103934665603ull;
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ull;
  }
  return h;
}

template <class T>
void put(std::string &k, const T &v) {
  k.append(reinterpret_cast<const char *>(&v), sizeof(T));
}

bool is_unary(int op) { return op >= CS_OP_NEG && op <= CS_OP_LOG; }
bool is_binary(int op) { return op >= CS_OP_ADD && op <= CS_OP_MAX; }
bool is_array(int op) { return op == CS_OP_PARAMETER || op == CS_OP_TRANSFORM || op == CS_OP_FILL; }

// A validated, canonically ordered DAG.
struct Dag {
  std::vector<cs_node> nodes;     // canonical order: operands before users
  std::vector<int> outputs;       // canonical ids
  std::vector<int64_t> out_shape;
  int nleaves = 0;
  std::vector<std::vector<int64_t>> leaf_shape;  // per leaf slot
};

// The array a leaf gather reads: the parameter plus the composed view matrix, or none
// for a bare parameter extract.
struct ArrayRef {
  int param = -1;            // canonical node id of the CS_OP_PARAMETER
  bool is_fill = false;      // array of a repeated float
  int fill_value = -1;       // float node id when is_fill
  bool has_view = false;
  int out_rank = 0;          // columns - 1 of matrix
  std::vector<double> matrix;  // rank_src x (out_rank + 1)
};

// NDimensionalAffineTransform.concatenate (NDimensionalAffineTransform.scala:37-95):
// m12 maps space 1 to space 2 (length2 rows, length1+1 columns), m01 maps space 0 to 1.
std::vector<double> compose(const std::vector<double> &m12, const std::vector<double> &m01,
                            int length2) {
  int length1 = static_cast<int>(m12.size()) / length2 - 1;
  int length0 = static_cast<int>(m01.size()) / length1 - 1;
  std::vector<double> m02(static_cast<size_t>((length0 + 1) * length2), 0.0);
  for (int i2 = 0; i2 < length2; ++i2) {
    for (int i0 = 0; i0 < length0; ++i0) {
      double acc = 0.0;
      for (int i1 = 0; i1 < length1; ++i1)
        acc += m12[i2 * (length1 + 1) + i1] * m01[i1 * (length0 + 1) + i0];
      m02[i2 * (length0 + 1) + i0] = acc;
    }
    double acc = m12[i2 * (length1 + 1) + length1];
    for (int i1 = 0; i1 < length1; ++i1)
      acc += m12[i2 * (length1 + 1) + i1] * m01[i1 * (length0 + 1) + length0];
    m02[i2 * (length0 + 1) + length0] = acc;
  }
  return m02;
}

// Validates the caller's nodes and renumbers the part reachable from the outputs in
// first-visit depth-first order (operand a before b). Leaf slots must already follow that
// order: it is the argument order the caller binds (Tensors.scala:230-251).
Dag canonicalize(const cs_node *nodes, int nnodes, const int32_t *outputs, int noutputs,
                 int out_rank, const int32_t *out_shape) {
  if (nnodes <= 0 || !nodes) fail(CS_INVALID_VALUE, "empty expression");
  if (noutputs <= 0 || !outputs) fail(CS_INVALID_VALUE, "no outputs");
  if (out_rank < 0 || out_rank > CS_MAX_RANK) fail(CS_INVALID_VALUE, "output rank out of range");
  if (out_rank > 0 && !out_shape) fail(CS_INVALID_VALUE, "out_shape is null");
  Dag dag;
  int64_t total = 1;
  for (int i = 0; i < out_rank; ++i) {
    if (out_shape[i] < 0) fail(CS_INVALID_VALUE, "negative output dimension");
    dag.out_shape.push_back(out_shape[i]);
    total *= out_shape[i];
  }
  // Index arithmetic is 32-bit, as in the reference (int indices); outputs stay below 2^31.
  if (total * noutputs >= (int64_t(1) << 31))
    fail(CS_INVALID_BUFFER_SIZE, "fused outputs are limited to 2^31 - 1 elements");

  std::vector<int> canon(static_cast<size_t>(nnodes), -1);
  int next_leaf = 0;
  // Value numbering: a node equal to an earlier canonical node (same op, canonical operands
  // and attributes) becomes that node, so structurally equal subtrees built as separate
  // objects (the same view taken once per term, as the convolution idiom does) are gathered
  // and computed once. The expressions are pure, so results are unchanged.
  std::unordered_map<std::string, int> seen;
  auto value_key = [&](const cs_node &c) {
    std::string k;
    put(k, c.op);
    put(k, c.a);
    put(k, c.b);
    if (c.op == CS_OP_PARAMETER || c.op == CS_OP_LITERAL) {
      uint32_t bits;
      std::memcpy(&bits, &c.value, 4);
      put(k, bits);
    }
    if (c.op == CS_OP_PARAMETER) put(k, c.param);
    if (c.op == CS_OP_PARAMETER || c.op == CS_OP_TRANSFORM) {
      put(k, c.rank);
      for (int i = 0; i < c.rank; ++i) put(k, c.shape[i]);
    }
    if (c.op == CS_OP_TRANSFORM) {
      const cs_node &opnd = dag.nodes[static_cast<size_t>(c.a)];
      const int rows = (opnd.op == CS_OP_PARAMETER || opnd.op == CS_OP_TRANSFORM) ? opnd.rank : 0;
      for (int i = 0; i < rows * (c.rank + 1); ++i) put(k, c.matrix[i]);
    }
    return k;
  };
  // Iterative DFS: pre-order for leaf numbering, post-order for emission.
  struct Frame { int node; int stage; };
  for (int oi = 0; oi < noutputs; ++oi) {
    int root = outputs[oi];
    if (root < 0 || root >= nnodes) fail(CS_INVALID_VALUE, "output index out of range");
    std::vector<Frame> stack{{root, 0}};
    std::vector<char> on_stack(static_cast<size_t>(nnodes), 0);
    while (!stack.empty()) {
      Frame &f = stack.back();
      int id = f.node;
      if (canon[id] >= 0) { stack.pop_back(); continue; }
      const cs_node &n = nodes[id];
      if (f.stage == 0) {
        if (on_stack[id]) fail(CS_INVALID_VALUE, "expression graph has a cycle");
        on_stack[id] = 1;
        if (n.op == CS_OP_PARAMETER) {
          if (n.param != next_leaf)
            fail(CS_INVALID_VALUE, "leaf slots must be numbered in first-visit order: expected " +
                                       std::to_string(next_leaf) + ", got " + std::to_string(n.param));
          if (n.rank < 0 || n.rank > CS_MAX_RANK) fail(CS_INVALID_VALUE, "leaf rank out of range");
          // Gather offsets are 32-bit sums of index * stride, like the reference's int
          // indices (and its Array[Float] leaves), so a leaf stays below 2^31 elements.
          int64_t leaf_elements = 1;
          for (int d = 0; d < n.rank; ++d) {
            if (n.shape[d] < 0) fail(CS_INVALID_VALUE, "negative leaf dimension");
            leaf_elements *= n.shape[d];
            if (leaf_elements >= (int64_t(1) << 31))
              fail(CS_INVALID_BUFFER_SIZE, "a leaf is limited to 2^31 - 1 elements");
          }
          next_leaf++;
        }
        f.stage = 1;
        auto want = [&](int child) {
          if (child < 0 || child >= nnodes) fail(CS_INVALID_VALUE, "operand index out of range");
          if (canon[child] < 0) stack.push_back({child, 0});
        };
        // Push b first so that a is visited first.
        bool two = is_binary(n.op);
        bool one = is_unary(n.op) || n.op == CS_OP_TRANSFORM || n.op == CS_OP_FILL || n.op == CS_OP_EXTRACT;
        if (!two && !one && n.op != CS_OP_LITERAL && n.op != CS_OP_PARAMETER)
          fail(CS_INVALID_VALUE, "unknown op " + std::to_string(n.op));
        if (two) { want(n.b); }
        if (two || one) { want(n.a); }
        // The vector may have grown: `f` is no longer valid here.
        continue;
      }
      // post-order
      cs_node c = n;
      auto map = [&](int child) { return canon[child]; };
      if (is_binary(n.op)) { c.a = map(n.a); c.b = map(n.b); }
      else if (is_unary(n.op) || n.op == CS_OP_TRANSFORM || n.op == CS_OP_FILL || n.op == CS_OP_EXTRACT) { c.a = map(n.a); c.b = -1; }
      else { c.a = c.b = -1; }
      const std::string vk = value_key(c);
      auto hit = seen.find(vk);
      if (hit != seen.end()) {
        canon[id] = hit->second;
      } else {
        canon[id] = static_cast<int>(dag.nodes.size());
        dag.nodes.push_back(c);
        seen.emplace(vk, canon[id]);
      }
      on_stack[id] = 0;
      stack.pop_back();
    }
    dag.outputs.push_back(canon[root]);
  }
  dag.nleaves = next_leaf;
  dag.leaf_shape.resize(static_cast<size_t>(next_leaf));
  // Type checks.
  for (auto &n : dag.nodes) {
    auto arr = [&](int i) { return is_array(dag.nodes[i].op); };
    switch (n.op) {
      case CS_OP_PARAMETER:
        dag.leaf_shape[n.param].assign(n.shape, n.shape + n.rank);
        for (int i = 0; i < n.rank; ++i)
          if (n.shape[i] < 0) fail(CS_INVALID_VALUE, "negative leaf dimension");
        break;
      case CS_OP_TRANSFORM:
        if (!arr(n.a)) fail(CS_INVALID_VALUE, "transform of a non-array");
        if (n.rank < 0 || n.rank > CS_MAX_RANK) fail(CS_INVALID_VALUE, "transform rank out of range");
        break;
      case CS_OP_FILL:
        if (arr(n.a)) fail(CS_INVALID_VALUE, "fill of an array");
        break;
      case CS_OP_EXTRACT:
        if (!arr(n.a)) fail(CS_INVALID_VALUE, "extract of a non-array");
        break;
      case CS_OP_LITERAL:
        break;
      default:
        if (arr(n.a) || (is_binary(n.op) && arr(n.b)))
          fail(CS_INVALID_VALUE, "arithmetic on an array");
    }
  }
  for (int o : dag.outputs)
    if (is_array(dag.nodes[o].op)) fail(CS_INVALID_VALUE, "an output must be a float term");
  return dag;
}

// Rank of the source array a TRANSFORM chain ends in; views of fills stay fills.
ArrayRef resolve_array(const Dag &dag, int id) {
  const cs_node &n = dag.nodes[id];
  if (n.op == CS_OP_PARAMETER) {
    ArrayRef r;
    r.param = id;
    return r;
  }
  if (n.op == CS_OP_FILL) {
    ArrayRef r;
    r.is_fill = true;
    r.fill_value = n.a;
    return r;
  }
  // TRANSFORM: ArrayParameter.transform creates a view (OpenCLKernelBuilder.scala:431-433);
  // a view's transform composes (OpenCLKernelBuilder.scala:334-340); a fill ignores it
  // (OpenCLKernelBuilder.scala:491-493).
  ArrayRef inner = resolve_array(dag, n.a);
  if (inner.is_fill) return inner;
  int src_rank = dag.nodes[inner.param].rank;
  int cols = n.rank + 1;
  int rows = inner.has_view ? inner.out_rank : src_rank;
  std::vector<double> m(n.matrix, n.matrix + static_cast<size_t>(rows) * cols);
  if (inner.has_view) {
    m = compose(inner.matrix, m, src_rank);
  }
  inner.has_view = true;
  inner.out_rank = n.rank;
  inner.matrix = m;
  return inner;
}

// java.text.DecimalFormat() with grouping off (OpenCLKernelBuilder.scala:14-32): at most
// three fraction digits, HALF_EVEN on the exact binary value, no trailing zeros.
std::string decimal_format(double v) {
  char buf[512];
  std::snprintf(buf, sizeof buf, "%.3f", v);  // glibc rounds the exact value, ties to even
  std::string s = buf;
  if (s.find('.') != std::string::npos) {
    while (!s.empty() && s.back() == '0') s.pop_back();
    if (!s.empty() && s.back() == '.') s.pop_back();
  }
  if (s == "-0") s = "-0";
  return s;
}

std::string float_literal(float v) {
  if (std::isnan(v)) return "__builtin_nanf(\"\")";
  if (std::isinf(v)) return v > 0 ? "__builtin_huge_valf()" : "(-__builtin_huge_valf())";
  char buf[64];
  std::snprintf(buf, sizeof buf, "%af", static_cast<double>(v));
  return buf;
}

struct Options {
  int math_mode = CS_MATH_PRECISE;
  int V = 0, U = 0, block = 0, nt = 0;
  int order = CS_ORDER_AUTO;
  bool sum = false;  // the kernel adds its single output into FP64 block partials instead of storing it
};

// A body of more than this many canonical nodes is large: it is compiled for its own workgroup
// size (launch_bound) and launched by default at one element per vector and per thread.
constexpr size_t kLargeBody = 256;

Options resolve_options(const cs_fused_options *o, const Dag &dag) {
  Options r;
  if (o) {
    r.math_mode = o->math_mode;
    r.V = o->vector_width;
    r.U = o->items_per_thread;
    r.block = o->block_size;
    r.nt = o->nontemporal_store;
    r.order = o->launch_order;
  }
  if (r.order != CS_ORDER_AUTO && r.order != CS_ORDER_LINEAR) fail(CS_INVALID_VALUE, "unknown launch order");
  if (r.math_mode != CS_MATH_PRECISE && r.math_mode != CS_MATH_FAST && r.math_mode != CS_MATH_ACCURATE)
    fail(CS_INVALID_VALUE, "unknown math mode");
  int64_t last = dag.out_shape.empty() ? 1 : dag.out_shape.back();
  int64_t total = 1;
  for (auto s : dag.out_shape) total *= s;
  // Defaults: 4 elements per vector, 2 vectors per thread from 2^22 elements. A large body
  // holds hundreds of values per element, so it defaults to 1 and 1 (the convolution idiom:
  // 9.3 us at V=1 against 25 us at V=2 and 50 us at V=4; two vectors per thread spill, 129 us).
  const bool large = dag.nodes.size() > kLargeBody;
  if (r.V == 0) r.V = large ? 1 : 4;
  if (r.V != 1 && r.V != 2 && r.V != 4 && r.V != 8) fail(CS_INVALID_VALUE, "vector_width must be 1, 2, 4 or 8");
  while (r.V > 1 && last % r.V != 0) r.V /= 2;
  if (r.U == 0) r.U = !large && total >= (int64_t(1) << 22) ? 2 : 1;
  if (r.U < 1 || r.U > 8) fail(CS_INVALID_VALUE, "items_per_thread must be 1..8");
  if (r.block == 0) r.block = 256;
  if (r.block < 64 || r.block > 1024 || r.block % 64 != 0)
    fail(CS_INVALID_WORK_GROUP_SIZE, "block_size must be a multiple of 64 in 64..1024");
  return r;
}

// Rows of a TRANSFORM matrix: the rank of the space its operand lives in.
int transform_rows(const Dag &dag, const cs_node &n) {
  const cs_node &opnd = dag.nodes[n.a];
  return (opnd.op == CS_OP_PARAMETER || opnd.op == CS_OP_TRANSFORM) ? opnd.rank : 0;
}

std::string canonical_key(const Dag &dag, const Options &o) {
  std::string k;
  k.reserve(dag.nodes.size() * 64 + 64);
  put(k, static_cast<int32_t>(dag.nodes.size()));
  for (const cs_node &n : dag.nodes) {
    put(k, n.op);
    put(k, n.a);
    put(k, n.b);
    if (n.op == CS_OP_PARAMETER || n.op == CS_OP_LITERAL) {
      uint32_t bits;
      std::memcpy(&bits, &n.value, 4);
      put(k, bits);
    }
    if (n.op == CS_OP_PARAMETER) put(k, n.param);
    if (n.op == CS_OP_PARAMETER || n.op == CS_OP_TRANSFORM) {
      put(k, n.rank);
      for (int i = 0; i < n.rank; ++i) put(k, n.shape[i]);
    }
    if (n.op == CS_OP_TRANSFORM) {
      int rows = transform_rows(dag, n);
      for (int i = 0; i < rows * (n.rank + 1); ++i) put(k, n.matrix[i]);
    }
  }
  put(k, static_cast<int32_t>(dag.outputs.size()));
  for (int out : dag.outputs) put(k, out);
  put(k, static_cast<int32_t>(dag.out_shape.size()));
  for (auto s : dag.out_shape) put(k, s);
  put(k, o.math_mode);
  put(k, o.V);
  put(k, o.U);
  put(k, o.block);
  put(k, o.nt);
  if (o.sum) put(k, int32_t(1));  // appended only in sum mode: element-wise keys are unchanged
  if (o.order != CS_ORDER_AUTO) put(k, int32_t(100 + o.order));  // likewise for the default order
  return k;
}

// The largest workgroup a compiled expression accepts. Small bodies are compiled for any
// size up to 1024 threads (so one compile serves every launch shape the autotuner tries);
// that caps a thread at 128 VGPRs, which a large body (a convolution's hundreds of terms)
// exceeds and spills. Those are compiled for their own workgroup size only, which gives them
// up to 512 registers.
int launch_bound(const Dag &dag, const Options &o) { return dag.nodes.size() > kLargeBody ? o.block : 1024; }

// ------------------------------------------------------------------------ code generation
// Kernel shape: each thread owns U vectors of V consecutive innermost outputs (vector u of
// thread t in block b is q = (b*U + u)*BLOCK + t, so a wave's accesses are contiguous).
// The body is emitted in phases -- coordinates for every u, every leaf gather for every u,
// then the arithmetic, then the stores -- so all of a thread's loads are in flight before
// its first use. A value is held as t<id>[U][V], or t<id>[U] when it is the same for the V
// elements of a vector (broadcast leaves, fills, and arithmetic on those).
class Emitter {
 public:
  Emitter(const Dag &dag, const Options &o) : dag_(dag), o_(o) {}

  // The whole kernel source: a stand-alone kernel over blockIdx.x.
  std::string generate(int64_t *vectors_out, const std::string &name, bool with_prelude = true) {
    std::ostringstream b, s;
    decls_.clear();
    prefix_ = name;
    body(b, "blockIdx.x", vectors_out);
    prefix_.clear();
    header(s, with_prelude);
    s << decls_;
    s << "extern \"C\" __global__ __launch_bounds__(" << launch_bound(dag_, o_) << ")" << waves_hint(*vectors_out)
      << " void " << name << "(";
    params(s);
    s << ") {\n" << b.str() << "}\n";
    return s.str();
  }

  // A large body (compiled for its own workgroup size) whose grid cannot fill the chip gets
  // the register budget of the occupancy that grid reaches, at least 2 waves per SIMD: with
  // 256 CUs of 4 SIMDs (MI355X), a grid of at most 2048 waves never has more than 2 waves on
  // a SIMD, and registers reserved for more only keep the compiler from holding loads in
  // flight (the convolution idiom's forward pass: 12.5 us without the hint, 9.9 us with it).
  static constexpr int64_t kComputeUnits = 256;
  std::string waves_hint(int64_t vectors) const {
    if (dag_.nodes.size() <= kLargeBody || o_.sum) return "";
    const int64_t block = o_.block, per_group = (block + 63) / 64;
    const int64_t groups = (vectors + block * o_.U - 1) / (block * o_.U);
    const int64_t per_simd = (groups * per_group + kComputeUnits * 4 - 1) / (kComputeUnits * 4);
    if (per_simd >= 8) return "";
    return " __attribute__((amdgpu_waves_per_eu(1, " + std::to_string(std::max<int64_t>(2, per_simd)) + ")))";
  }

  // The same body as a device function over a run of tiles [tile_begin, tile_end), for
  // batches: one call per work item, the loop inside. It is not inlined: inlined into one
  // switch, hundreds of bodies share values hoisted across the cases and the kernel's
  // register count (and occupancy) is set by the sum, not the largest. Pointers carry the
  // global address space (a function parameter is otherwise generic and every access a flat
  // one) and the workgroup size is the batch's constant.
  std::string device_function(const std::string &name, int block, int64_t *vectors_out) {
    gq_ = "CSMI_G ";
    bdim_ = std::to_string(block) + "u";
    std::ostringstream s;
    s << "__device__ __noinline__ void " << name << "(const unsigned tile_begin, const unsigned tile_end, ";
    params(s);
    s << ") {\n  for (unsigned tile = tile_begin; tile < tile_end; ++tile) {\n";
    body(s, "tile", vectors_out);
    s << "  }\n}\n";
    gq_.clear();
    bdim_ = "blockDim.x";
    return s.str();
  }

  // Typedefs (and the accurate math functions) the bodies need, once per module.
  static std::string prelude(int math_mode) {
    std::string p;
    p += std::string(kPreciseMathSource) + "\n";
    p += "typedef float csmi_f4 __attribute__((ext_vector_type(4), aligned(4)));\n";
    p += "typedef float csmi_f2 __attribute__((ext_vector_type(2), aligned(4)));\n";
    p += "typedef float csmi_v4 __attribute__((ext_vector_type(4)));\n";
    p += "#define CSMI_G __attribute__((address_space(1)))\n";
    // s + c += x with the rounding error of the addition kept in c (TwoSum); the steps are
    // exact only without contraction or reassociation, whatever the math mode
    p += "__device__ __forceinline__ void csmi_acc(float &s, float &c, float x) {\n"
         "#pragma clang fp contract(off) reassociate(off)\n"
         "  const float t = s + x;\n  const float bp = t - s;\n"
         "  c += (s - (t - bp)) + (x - bp);\n  s = t;\n}\n";
    return p;
  }

  int nleaves() const { return dag_.nleaves; }

 private:
  void header(std::ostringstream &s, bool with_prelude) {
    const int r = static_cast<int>(dag_.out_shape.size());
    s << "// Generated by csmi for gfx950: " << dag_.nodes.size() << " nodes, " << dag_.nleaves
      << " leaves, output [";
    for (int i = 0; i < r; ++i) s << (i ? "," : "") << dag_.out_shape[i];
    s << "] x " << dag_.outputs.size() << ", V=" << o_.V << " U=" << o_.U << "\n";
    if (with_prelude) s << prelude(o_.math_mode);
  }

  void params(std::ostringstream &s) {
    for (int i = 0; i < dag_.nleaves; ++i) s << "const " << gq_ << "float* __restrict__ p" << i << ", ";
    if (o_.sum)
      s << "double* __restrict__ partials";
    else
      s << gq_ << "float* __restrict__ out";
  }

  // tile: the index of the block's slice of blockDim.x * U vectors.
  // Launch order. A leaf read through a view whose source rows run along an outer output
  // dimension d (the source's innermost coordinate moves with d alone, by +-1: a transpose)
  // is read one element per source row by a wave that walks the output's innermost
  // dimension, so each load takes a few bytes of a line and the rest must wait in the caches.
  // Then workgroups take tiles of R positions of d by C vectors of the dimensions after d
  // (C fastest): R consecutive elements of each source row are read together, and the output
  // and every other leaf are still written and read in runs of C vectors (whole rows when
  // they are short). d is the dimension with the most such leaf elements. R is d's extent up
  // to 32, else the size in 16..32 that pads it least; C likewise up to 16, in 8..16. The
  // tiles cover the padded extents, whose extra positions are idle lanes; the order stays
  // linear when R < 4 or the padding exceeds 1/8 of the output. Which thread computes which
  // element changes, not how: results are the same bits.
  struct Order {
    bool tiled = false;
    int64_t Sd = 0, Bv = 0, R = 0, C = 0, Sdp = 0, Bvp = 0, iter = 0;
  };
  static int64_t least_padding(int64_t extent, int64_t lo, int64_t hi) {
    if (extent <= hi) return extent;
    int64_t best = hi, pad = (extent + hi - 1) / hi * hi;
    for (int64_t c = hi - 1; c >= lo; --c) {
      const int64_t p = (extent + c - 1) / c * c;
      if (p < pad) pad = p, best = c;
    }
    return best;
  }
  Order choose_order(int64_t vectors) const {
    Order ord;
    const int r = static_cast<int>(dag_.out_shape.size());
    if (o_.order == CS_ORDER_LINEAR || o_.sum || r < 2) return ord;
    const int64_t V = o_.V;
    if (dag_.out_shape[r - 1] % V != 0) return ord;
    std::vector<double> weight(static_cast<size_t>(r), 0.0);
    for (const cs_node &n : dag_.nodes) {
      if (n.op != CS_OP_EXTRACT) continue;
      ArrayRef ar = resolve_array(dag_, n.a);
      if (ar.is_fill || !ar.has_view || ar.out_rank != r) continue;
      const cs_node &p = dag_.nodes[ar.param];
      if (p.rank < 2 || p.shape[p.rank - 1] < 2) continue;
      const int cols = r + 1, y = p.rank - 1;
      int drive = -1, count = 0;
      bool unit = true;
      for (int x = 0; x < r; ++x) {
        const double m = ar.matrix[static_cast<size_t>(y) * cols + x];
        if (m == 0.0) continue;
        ++count;
        drive = x;
        unit = unit && (m == 1.0 || m == -1.0);
      }
      if (count != 1 || !unit || drive == r - 1) continue;
      double elements = 1.0;
      for (int i = 0; i < p.rank; ++i) elements *= p.shape[i];
      weight[static_cast<size_t>(drive)] += elements;
    }
    int d = -1;
    for (int x = 0; x < r - 1; ++x)
      if (weight[static_cast<size_t>(x)] > 0.0 && (d < 0 || weight[static_cast<size_t>(x)] > weight[static_cast<size_t>(d)]))
        d = x;
    if (d < 0) return ord;
    int64_t B = 1;
    for (int x = d + 1; x < r; ++x) B *= dag_.out_shape[static_cast<size_t>(x)];
    const int64_t Sd = dag_.out_shape[static_cast<size_t>(d)], Bv = B / V;
    const int64_t R = least_padding(Sd, 16, 32), C = least_padding(Bv, 8, 16);
    if (R < 4) return ord;
    const int64_t Sdp = (Sd + R - 1) / R * R, Bvp = (Bv + C - 1) / C * C;
    const int64_t iter = vectors / (Sd * Bv) * Sdp * Bvp;
    if (iter * 8 > vectors * 9 || iter >= (int64_t(1) << 31)) return ord;
    ord.tiled = true;
    ord.Sd = Sd, ord.Bv = Bv, ord.R = R, ord.C = C, ord.Sdp = Sdp, ord.Bvp = Bvp, ord.iter = iter;
    return ord;
  }

  void body(std::ostringstream &s, const std::string &tile, int64_t *vectors_out) {
    const int r = static_cast<int>(dag_.out_shape.size());
    const int V = o_.V, U = o_.U;
    int64_t total = 1;
    for (auto d : dag_.out_shape) total *= d;
    const int64_t vectors = total / V;
    const int nout = static_cast<int>(dag_.outputs.size());
    uniform_.clear();
    fill_of_.clear();
    chunks_.clear();
    const Order ord = choose_order(vectors);
    // the launch covers the tiles' padded extents when tiled
    *vectors_out = ord.tiled ? ord.iter : vectors;
    s << "  const unsigned vbase = (" << tile << ") * (" << bdim_ << " * " << U << "u) + threadIdx.x;\n";
    // phase 0: vector index, validity and coordinates of element 0 of every vector
    s << "  unsigned q[" << U << "]; bool live[" << U << "];\n";
    for (int d = 0; d < r; ++d) s << "  int c" << d << "[" << U << "];\n";
    s << "#pragma unroll\n  for (int u = 0; u < " << U << "; ++u) {\n";
    s << "    q[u] = vbase + u * " << bdim_ << ";\n";
    if (!ord.tiled) {
      s << "    live[u] = q[u] < " << vectors << "u;\n";
    } else {
      // iteration index -> vector index: tiles of R positions of dimension d by C vectors of
      // the dimensions after it, C fastest, over the padded extents; tiles in row-major order
      // of (outer, d, inner); a padding position is an idle lane
      const int64_t T = ord.R * ord.C, tiles_b = ord.Bvp / ord.C, tiles_d = ord.Sdp / ord.R;
      s << "    live[u] = q[u] < " << ord.iter << "u;\n";
      s << "    {\n";
      s << "      const unsigned t = q[u] % " << T << "u, tl = q[u] / " << T << "u;\n";
      s << "      const unsigned bt = tl % " << tiles_b << "u, rt = tl / " << tiles_b << "u;\n";
      s << "      const unsigned dt = rt % " << tiles_d << "u, ot = rt / " << tiles_d << "u;\n";
      s << "      const unsigned dd = dt * " << ord.R << "u + t / " << ord.C << "u, bb = bt * " << ord.C << "u + t % "
        << ord.C << "u;\n";
      s << "      live[u] = live[u] && dd < " << ord.Sd << "u && bb < " << ord.Bv << "u;\n";
      s << "      q[u] = live[u] ? (ot * " << ord.Sd << "u + dd) * " << ord.Bv << "u + bb : " << vectors << "u;\n";
      s << "    }\n";
    }
    s << "    unsigned rest = q[u] * " << V << "u;\n";
    for (int d = r - 1; d >= 0; --d) {
      if (d > 0)
        s << "    c" << d << "[u] = (int)(rest % " << dag_.out_shape[d] << "u); rest /= " << dag_.out_shape[d] << "u;\n";
      else
        s << "    c0[u] = (int)rest;\n";
    }
    s << "  }\n";
    body_ = &s;
    staged_.clear();
    if (gq_.empty()) stage_constants(s, prefix_);
    plan_row_groups();
    // phase 1: gathers
    for (size_t i = 0; i < dag_.nodes.size(); ++i)
      if (dag_.nodes[i].op == CS_OP_EXTRACT) emit_gather(static_cast<int>(i));
    // phase 2: arithmetic
    for (size_t i = 0; i < dag_.nodes.size(); ++i) {
      int op = dag_.nodes[i].op;
      if (op != CS_OP_EXTRACT && !is_array(op) && op != CS_OP_LITERAL) emit_arith(static_cast<int>(i));
    }
    if (o_.sum) {
      emit_block_sum(s);
      return;
    }
    // phase 3: stores -- out[(e0 + v) * nout + j] is one contiguous run of V * nout floats
    const int run = V * nout;
    s << "#pragma unroll\n  for (int u = 0; u < " << U << "; ++u) {\n";
    s << "    if (!live[u]) continue;\n";
    s << "    " << gq_ << "float* __restrict__ dst = out + (unsigned long long)q[u] * " << run << "u;\n";
    std::vector<std::string> vals;
    for (int v = 0; v < V; ++v)
      for (int j = 0; j < nout; ++j) vals.push_back(value_of(dag_.outputs[j], v));
    int i = 0;
    while (i < run) {
      if (run % 4 == 0 && run - i >= 4) {
        s << "    " << store("dst + " + std::to_string(i),
                             "csmi_v4{" + vals[i] + ", " + vals[i + 1] + ", " + vals[i + 2] + ", " + vals[i + 3] + "}",
                             "csmi_v4", run)
          << "\n";
        i += 4;
      } else {
        s << "    " << store("dst + " + std::to_string(i), vals[i], "float", run) << "\n";
        i += 1;
      }
    }
    s << "  }\n";
  }

  // Sum mode, in place of the stores: each thread adds its elements with a compensation term
  // (csmi_acc), the (sum, compensation) pairs become FP64 and are added by a fixed xor-shuffle
  // tree in each wave and in wave order across the block, and thread 0 writes the block's
  // partial. The element values are the ones the stores would write.
  void emit_block_sum(std::ostringstream &s) {
    const int V = o_.V, U = o_.U;
    s << "  float csmi_s = 0.0f, csmi_c = 0.0f;\n";
    s << "#pragma unroll\n  for (int u = 0; u < " << U << "; ++u) {\n    if (!live[u]) continue;\n";
    for (int v = 0; v < V; ++v) s << "    csmi_acc(csmi_s, csmi_c, " << value_of(dag_.outputs[0], v) << ");\n";
    s << "  }\n";
    // c only while s is finite: an infinite or NaN sum stays so, and its error terms are NaN
    s << "  double csmi_d = __builtin_isfinite(csmi_s) ? (double)csmi_s + (double)csmi_c : (double)csmi_s;\n";
    s << "#pragma unroll\n  for (int off = 32; off >= 1; off >>= 1) csmi_d += __shfl_xor(csmi_d, off, 64);\n";
    s << "  __shared__ double csmi_lw[16];\n";
    s << "  if ((threadIdx.x & 63u) == 0u) csmi_lw[threadIdx.x >> 6] = csmi_d;\n";
    s << "  __syncthreads();\n";
    s << "  if (threadIdx.x == 0u) {\n    double t = 0.0;\n";
    s << "    for (unsigned w = 0; w < " << bdim_ << " / 64u; ++w) t += csmi_lw[w];\n";
    s << "    partials[" << (bdim_ == "blockDim.x" ? "blockIdx.x" : "tile") << "] = t;\n  }\n";
  }

  // Non-temporal stores. When one store instruction covers the wave's lanes contiguously (a
  // run of 1 or 4 floats per vector) they also write through at device scope (`sc1 nt`, a
  // vector store): the output is not read again, and no dirty line is left in L2 for the
  // release at the end of the kernel, which otherwise writes it back between this launch and
  // the next. Longer or odd runs leave each instruction every other piece of a line, which
  // written through becomes partial-line writes: those stay plain non-temporal stores.
  std::string store(const std::string &addr, const std::string &val, const std::string &type, int run) {
    if (o_.nt) {
      const bool whole = (run == 4 && type == "csmi_v4") || (run == 1 && type == "float");
      if (!whole) return "__builtin_nontemporal_store(" + val + ", (" + gq_ + type + "*)(" + addr + "));";
      const std::string op = type == "csmi_v4" ? "global_store_dwordx4" : "global_store_dword";
      return "asm volatile(\"" + op + " %0, %1, off sc1 nt\" :: \"v\"((" + gq_ + type + "*)(" + addr +
             ")), \"v\"(" + val + ") : \"memory\");";
    }
    return "*(" + gq_ + type + "*)(" + addr + ") = " + val + ";";
  }

  // True when node `id` cannot exceed 88 (so exp of it cannot overflow): the negation of a
  // value that is never negative (a square x*x of one node, or an absolute value). NaN
  // inputs stay NaN either way.
  bool bounded_above(int id) const {
    const cs_node &n = dag_.nodes[static_cast<size_t>(id)];
    if (n.op != CS_OP_NEG) return false;
    const cs_node &m = dag_.nodes[static_cast<size_t>(n.a)];
    return (m.op == CS_OP_MUL && m.a == m.b) || m.op == CS_OP_ABS;
  }

  // extract(fill(x)) stands for x
  int resolve(int id) const {
    auto it = fill_of_.find(id);
    return it == fill_of_.end() ? id : resolve(it->second);
  }

  // FP32 value of float node `id` for vector u (the loop variable) and element v.
  std::string value_of(int id, int v) {
    id = resolve(id);
    const cs_node &n = dag_.nodes[id];
    if (n.op == CS_OP_LITERAL) return float_literal(n.value);
    if (uniform_.count(id)) return "t" + std::to_string(id) + "[u]";
    return "t" + std::to_string(id) + "[u][" + std::to_string(v) + "]";
  }
  std::string value_in_loop(int id) {
    id = resolve(id);
    const cs_node &n = dag_.nodes[id];
    if (n.op == CS_OP_LITERAL) return float_literal(n.value);
    if (uniform_.count(id)) return "t" + std::to_string(id) + "[u]";
    return "t" + std::to_string(id) + "[u][v]";
  }
  bool is_uniform(int id) const {
    id = resolve(id);
    const cs_node &n = dag_.nodes[id];
    return n.op == CS_OP_LITERAL || uniform_.count(id) > 0;
  }

  // One SSA statement per operation and element (OpenCLKernelBuilder.scala:516-571,
  // 271-325). A value that is the same for all V elements is computed once per vector.
  void emit_arith(int id) {
    const cs_node &n = dag_.nodes[id];
    std::ostringstream &s = *body_;
    const int V = o_.V, U = o_.U;
    bool uni = is_uniform(n.a) && (!is_binary(n.op) || is_uniform(n.b));
    std::string name = "t" + std::to_string(id);
    if (uni) {
      uniform_.insert(id);
      s << "  float " << name << "[" << U << "];\n";
    } else {
      s << "  float " << name << "[" << U << "][" << V << "];\n";
    }
    s << "#pragma unroll\n  for (int u = 0; u < " << U << "; ++u)\n";
    s << "#pragma unroll\n  for (int v = 0; v < " << (uni ? 1 : V) << "; ++v) ";
    std::string lhs = uni ? name + "[u]" : name + "[u][v]";
    std::string a = value_in_loop(n.a), b = is_binary(n.op) ? value_in_loop(n.b) : "";
    const bool fast = o_.math_mode == CS_MATH_FAST, acc = o_.math_mode == CS_MATH_ACCURATE;
    std::string e;
    switch (n.op) {
      case CS_OP_NEG: e = "-" + a; break;
      case CS_OP_ABS: e = "fabsf(" + a + ")"; break;
      case CS_OP_SQRT: e = "sqrtf(" + a + ")"; break;
      case CS_OP_TANH: e = (acc ? "csmi_tanhf(" : "csmi_tanhf_p(") + a + ")"; break;
      case CS_OP_EXP:
        e = (fast ? "__expf(" : acc ? "csmi_expf(" : bounded_above(n.a) ? "csmi_expf_np(" : "expf(") + a + ")";
        break;
      case CS_OP_LOG: e = (fast ? "__logf(" : "logf(") + a + ")"; break;
      case CS_OP_ADD: e = a + " + " + b; break;
      case CS_OP_SUB: e = a + " - " + b; break;
      case CS_OP_MUL: e = a + " * " + b; break;
      case CS_OP_DIV: e = a + " / " + b; break;
      // OpenCL C has no float `%` (the reference's text fails to build); the C fmod is the
      // meaning the operator has everywhere else.
      case CS_OP_MOD: e = "fmodf(" + a + ", " + b + ")"; break;
      case CS_OP_MIN: e = "fminf(" + a + ", " + b + ")"; break;
      case CS_OP_MAX: e = "fmaxf(" + a + ", " + b + ")"; break;
      default: fail(CS_INVALID_VALUE, "unexpected op in codegen");
    }
    s << lhs << " = " << e << ";\n";
  }

  // A leaf read at the current output coordinates (OpenCLKernelBuilder.scala:348-411 for
  // views, 435-455 for a bare parameter).
  struct Row {
    bool empty = true;        // all-zero row: index 0, no bounds check
    bool fractional = false;  // double arithmetic, truncated by (int)
    std::string expr;         // index of element 0 of the vector
    int64_t inner_step = 0;   // integer coefficient of the innermost output coordinate
    bool upper_only = false;  // bare parameter: only idx < dim is checked
    int64_t shift = 0;        // integral constant term
    bool others_by4 = true;   // integral coefficients of the other coordinates, all multiples of 4
    std::string unshifted;    // expr without the integral constant term
  };

  // One Row per source dimension of the leaf a gather reads: the index expression of element
  // 0 of the vector and its properties.
  std::vector<Row> gather_rows(const ArrayRef &ar, const cs_node &p) const {
    const int r = static_cast<int>(dag_.out_shape.size());
    const int src_rank = p.rank;
    std::vector<Row> rows(static_cast<size_t>(src_rank));
    if (!ar.has_view) {
      if (src_rank > r)
        fail(CS_INVALID_VALUE, "a leaf of rank " + std::to_string(src_rank) +
                                   " cannot be read directly in a rank-" + std::to_string(r) + " kernel");
      for (int y = 0; y < src_rank; ++y) {
        rows[y].empty = false;
        rows[y].upper_only = true;
        rows[y].expr = "c" + std::to_string(y) + "[u]";
        rows[y].inner_step = (y == r - 1) ? 1 : 0;
      }
    } else {
      if (ar.out_rank != r)
        fail(CS_INVALID_VALUE, "a view of rank " + std::to_string(ar.out_rank) + " is read in a rank-" +
                                   std::to_string(r) + " kernel");
      const int cols = r + 1;
      for (int y = 0; y < src_rank; ++y) {
        Row &row = rows[y];
        std::vector<std::string> terms;
        bool frac = false;
        for (int x = 0; x < cols; ++x) {
          double m = ar.matrix[static_cast<size_t>(y) * cols + x];
          if (m == 0.0) continue;
          std::string lit = decimal_format(m);
          if (lit.find('.') != std::string::npos) frac = true;
          if (x < cols - 1) {
            std::string cx = "c" + std::to_string(x) + "[u]";
            terms.push_back(m == 1.0 ? cx : cx + " * " + lit);
            if (x == r - 1 && lit.find('.') == std::string::npos) row.inner_step = std::stoll(lit);
            if (x != r - 1 && (lit.find('.') != std::string::npos || std::stoll(lit) % 4 != 0)) row.others_by4 = false;
          } else {
            terms.push_back(lit);
            if (lit.find('.') == std::string::npos) row.shift = std::stoll(lit);
          }
        }
        if (terms.empty()) continue;
        row.empty = false;
        row.fractional = frac;
        std::string e, u;
        for (size_t t = 0; t < terms.size(); ++t) e += (t ? " + " : "") + terms[t];
        const size_t nu = terms.size() - (row.shift != 0 ? 1 : 0);  // the constant term is last
        for (size_t t = 0; t < nu; ++t) u += (t ? " + " : "") + terms[t];
        row.expr = e;
        row.unshifted = u;
      }
    }
    return rows;
  }

  // The flat offset of a gather whose every index is an integral constant inside the leaf (an
  // element picked out by split, then broadcast), with its leaf in *leaf; -1 for any other
  // gather (one that moves with the coordinates, truncates a fraction, or reads the padding).
  int64_t constant_offset(int id, int *leaf) const {
    ArrayRef ar = resolve_array(dag_, dag_.nodes[id].a);
    if (ar.is_fill) return -1;
    const cs_node &p = dag_.nodes[ar.param];
    const int r = static_cast<int>(dag_.out_shape.size()), cols = r + 1;
    if (p.rank > 0 && (!ar.has_view || ar.out_rank != r)) return -1;
    int64_t off = 0, stride = 1;
    for (int y = p.rank - 1; y >= 0; --y) {
      for (int x = 0; x < r; ++x)
        if (ar.matrix[static_cast<size_t>(y) * cols + x] != 0.0) return -1;
      const double m = ar.matrix[static_cast<size_t>(y) * cols + r];
      int64_t i = 0;
      if (m != 0.0) {
        const std::string lit = decimal_format(m);
        if (lit.find('.') != std::string::npos) return -1;
        i = std::stoll(lit);
      }
      if (i < 0 || i >= p.shape[y]) return -1;
      off += i * stride;
      stride *= p.shape[y];
    }
    *leaf = p.param;
    return off;
  }

  // Constant tables in LDS. A body that reads many single elements of small leaves (the
  // weights and biases of the convolution idiom: 576 + 8 scalars, each broadcast over the
  // output) would otherwise issue one per-lane global load per scalar in every thread. Those
  // elements are copied once per workgroup into csmi_k, in the order the gathers read them
  // (so that consecutive reads are adjacent and load as one wide LDS read), through a table of
  // (slot, offset) pairs per leaf kept in constant memory; each gather then reads its slot (a
  // broadcast LDS read). The values are the leaf's own, so results are the same bits. Kernels
  // only (a batch's bodies would add up their tables), and only with at least kMinStaged such
  // gathers, in at most kMaxStaged floats.
  static constexpr int64_t kMinStaged = 32, kMaxStaged = 4096;
  void stage_constants(std::ostringstream &s, const std::string &prefix) {
    std::vector<std::pair<int, int64_t>> reads;  // (leaf, offset), in node order
    std::map<int, std::set<int64_t>> distinct;
    for (size_t i = 0; i < dag_.nodes.size(); ++i) {
      if (dag_.nodes[i].op != CS_OP_EXTRACT) continue;
      int leaf = -1;
      const int64_t off = constant_offset(static_cast<int>(i), &leaf);
      if (off < 0) continue;
      reads.push_back({leaf, off});
      distinct[leaf].insert(off);
    }
    std::set<int> chosen;
    int64_t total = 0;
    for (const auto &kv : distinct) {
      const int64_t n = static_cast<int64_t>(kv.second.size());
      if (total + n > kMaxStaged) continue;
      chosen.insert(kv.first);
      total += n;
    }
    if (total < kMinStaged) return;
    std::map<int, std::vector<std::pair<int64_t, int64_t>>> table;  // leaf -> (slot, offset)
    for (const auto &rd : reads) {
      if (!chosen.count(rd.first) || staged_.count(rd)) continue;
      const int64_t slot = static_cast<int64_t>(staged_.size());
      staged_[rd] = slot;
      table[rd.first].push_back({slot, rd.second});
    }
    std::ostringstream d;
    for (const auto &kv : table) {
      for (int half = 0; half < 2; ++half) {
        d << "__constant__ unsigned " << prefix << (half ? "_ko" : "_ks") << kv.first << "[" << kv.second.size()
          << "] = {";
        for (size_t i = 0; i < kv.second.size(); ++i)
          d << (i ? (i % 16 ? ", " : ",\n  ") : "") << (half ? kv.second[i].second : kv.second[i].first);
        d << "};\n";
      }
    }
    decls_ += d.str();
    s << "  __shared__ float csmi_k[" << total << "];\n";
    for (const auto &kv : table)
      s << "  for (unsigned i = threadIdx.x; i < " << kv.second.size() << "u; i += blockDim.x) csmi_k[" << prefix
        << "_ks" << kv.first << "[i]] = p" << kv.first << "[" << prefix << "_ko" << kv.first << "[i]];\n";
    s << "  __syncthreads();\n";
  }

  // Rows read together. With one element per vector (V = 1), gathers of one leaf whose
  // indices agree in every source dimension but the last, where each has an in-range
  // constant (the channels of a pixel, each split out and read through the same view, as in
  // the convolution idiom), read one source row: they share one bounds check and row address,
  // and the constants that fall two or more into one aligned 4-element chunk of the row load
  // it as one 16-byte access (the row length a multiple of 4). Each value is the element, or
  // the padding, that its own gather reads.
  struct RowGroup {
    std::vector<std::pair<int, int64_t>> members;  // (node, constant), in node order
    bool emitted = false;
  };
  void plan_row_groups() {
    row_groups_.clear();
    in_group_.clear();
    if (o_.V != 1) return;
    const int r = static_cast<int>(dag_.out_shape.size());
    std::map<std::string, RowGroup> cand;
    for (size_t i = 0; i < dag_.nodes.size(); ++i) {
      if (dag_.nodes[i].op != CS_OP_EXTRACT) continue;
      ArrayRef ar = resolve_array(dag_, dag_.nodes[i].a);
      if (ar.is_fill || !ar.has_view || ar.out_rank != r) continue;
      const cs_node &p = dag_.nodes[ar.param];
      if (p.rank < 2) continue;
      int leaf = -1;
      const int64_t off = constant_offset(static_cast<int>(i), &leaf);
      if (off >= 0 && staged_.count({leaf, off})) continue;
      const std::vector<Row> rows = gather_rows(ar, p);
      const Row &last = rows[static_cast<size_t>(p.rank - 1)];
      if (!last.empty && (last.fractional || !last.unshifted.empty())) continue;
      const int64_t k = last.empty ? 0 : last.shift;
      if (k < 0 || k >= p.shape[p.rank - 1]) continue;
      std::string key = "p" + std::to_string(p.param);
      bool ok = true;
      for (int y = 0; y + 1 < p.rank; ++y) {
        ok = ok && !rows[static_cast<size_t>(y)].fractional;
        key += "|" + (rows[static_cast<size_t>(y)].empty ? std::string("_") : rows[static_cast<size_t>(y)].expr);
      }
      if (ok) cand[key].members.push_back({static_cast<int>(i), k});
    }
    for (auto &kv : cand) {
      if (kv.second.members.size() < 2) continue;
      for (const auto &m : kv.second.members) in_group_[m.first] = kv.first;
      row_groups_[kv.first] = kv.second;
    }
  }

  // Emits every gather of a row group, at its first member.
  void emit_row_group(RowGroup &grp, const ArrayRef &ar) {
    std::ostringstream &s = *body_;
    const int U = o_.U;
    const cs_node &p = dag_.nodes[ar.param];
    const std::string ptr = "p" + std::to_string(p.param), pad = float_literal(p.value);
    const std::string g = "r" + std::to_string(grp.members.front().first);
    const std::vector<Row> rows = gather_rows(ar, p);
    const int L = p.rank - 1;
    int64_t elements = 1;
    for (int y = 0; y < p.rank; ++y) elements *= p.shape[y];
    std::map<int64_t, int> per_chunk;
    for (const auto &m : grp.members) ++per_chunk[m.second / 4];
    const bool chunks = p.shape[L] % 4 == 0 && elements > 0;
    for (const auto &m : grp.members) {
      s << "  float t" << m.first << "[" << U << "];\n";
      uniform_.insert(m.first);
    }
    s << "#pragma unroll\n  for (int u = 0; u < " << U << "; ++u) {\n";
    std::vector<std::string> conds{"live[u]"};
    std::string off;
    int64_t stride = p.shape[L];
    std::vector<std::string> terms;
    for (int y = L - 1; y >= 0; --y) {
      const Row &row = rows[static_cast<size_t>(y)];
      const std::string ix = g + "_i" + std::to_string(y);
      if (row.empty) {
        stride *= p.shape[y];
        continue;
      }
      s << "      const int " << ix << " = (int)(" << row.expr << ");\n";
      conds.push_back(ix + " >= 0");
      conds.push_back(ix + " < " + std::to_string(p.shape[y]));
      terms.push_back(ix + " * " + std::to_string(stride));
      stride *= p.shape[y];
    }
    for (auto it = terms.rbegin(); it != terms.rend(); ++it) off += (off.empty() ? "" : " + ") + *it;
    std::string cond;
    for (size_t c = 0; c < conds.size(); ++c) cond += (c ? " && " : "") + conds[c];
    s << "      const bool " << g << "_ok = " << cond << ";\n";
    s << "      const " << gq_ << "float* " << g << "_p = " << ptr << (off.empty() ? "" : " + (" + off + ")") << ";\n";
    std::set<int64_t> loaded;
    for (const auto &m : grp.members) {
      const int64_t q = m.second / 4;
      if (chunks && per_chunk[q] >= 2) {
        const std::string w = g + "_c" + std::to_string(q);
        if (loaded.insert(q).second)
          s << "      const csmi_f4 " << w << " = *(const " << gq_ << "csmi_f4*)(" << g << "_ok ? " << g << "_p + "
            << 4 * q << " : " << ptr << ");\n";
        s << "      t" << m.first << "[u] = " << g << "_ok ? " << w << "[" << m.second % 4 << "] : " << pad << ";\n";
      } else {
        s << "      t" << m.first << "[u] = " << g << "_ok ? " << g << "_p[" << m.second << "] : " << pad << ";\n";
      }
    }
    s << "  }\n";
    grp.emitted = true;
  }

  void emit_gather(int id) {
    std::ostringstream &s = *body_;
    const int V = o_.V, U = o_.U;
    const int r = static_cast<int>(dag_.out_shape.size());
    const cs_node &n = dag_.nodes[id];
    ArrayRef ar = resolve_array(dag_, n.a);
    std::string name = "t" + std::to_string(id);
    if (ar.is_fill) {
      // extract(fill(x)) == x; x is computed in the arithmetic phase, so alias it there
      fill_of_[id] = ar.fill_value;
      return;
    }
    if (!staged_.empty()) {
      int leaf = -1;
      const int64_t off = constant_offset(id, &leaf);
      auto it = off < 0 ? staged_.end() : staged_.find({leaf, off});
      if (it != staged_.end()) {
        uniform_.insert(id);
        s << "  float " << name << "[" << U << "];\n";
        s << "#pragma unroll\n  for (int u = 0; u < " << U << "; ++u) " << name << "[u] = live[u] ? csmi_k["
          << it->second << "] : " << float_literal(dag_.nodes[ar.param].value) << ";\n";
        return;
      }
    }
    auto member = in_group_.find(id);
    if (member != in_group_.end()) {
      RowGroup &grp = row_groups_[member->second];
      if (!grp.emitted) emit_row_group(grp, ar);
      return;
    }
    const cs_node &p = dag_.nodes[ar.param];
    const int src_rank = p.rank;
    const std::string ptr = "p" + std::to_string(p.param);
    const std::string pad = float_literal(p.value);
    if (src_rank == 0) {
      // `*x`: a scalar source is read without bounds (OpenCLKernelBuilder.scala:351-356).
      uniform_.insert(id);
      s << "  float " << name << "[" << U << "];\n";
      s << "#pragma unroll\n  for (int u = 0; u < " << U << "; ++u) " << name << "[u] = " << ptr << "[0];\n";
      return;
    }
    std::vector<Row> rows = gather_rows(ar, p);
    std::vector<int64_t> stride(static_cast<size_t>(src_rank), 1);
    for (int y = src_rank - 2; y >= 0; --y) stride[y] = stride[y + 1] * p.shape[y + 1];

    bool any_frac = false;
    int moving = -1, nmoving = 0;
    for (int y = 0; y < src_rank; ++y) {
      if (rows[y].empty) continue;
      any_frac = any_frac || rows[y].fractional;
      if (!rows[y].fractional && rows[y].inner_step != 0) {
        moving = y;
        ++nmoving;
      }
    }
    const bool broadcast = !any_frac && nmoving == 0;
    const bool contiguous = !any_frac && nmoving == 1 && moving == src_rank - 1 && rows[moving].inner_step == 1 && V > 1;
    const std::string g = "g" + std::to_string(id);
    // Per-row indices (element 0, or per element with the innermost coordinate advanced)
    auto emit_rows = [&](const std::string &cinner) {
      std::vector<std::string> conds;
      std::string off;
      for (int y = 0; y < src_rank; ++y) {
        const Row &row = rows[y];
        std::string ix = g + "_i" + std::to_string(y);
        if (row.empty) {
          s << "      const int " << ix << " = 0;\n";
        } else {
          std::string e = row.expr;
          if (!cinner.empty() && r > 0) e = replace_coord(e, "c" + std::to_string(r - 1) + "[u]", cinner);
          s << "      const int " << ix << " = (int)(" << e << ");\n";
          if (row.upper_only) {
            conds.push_back(ix + " < " + std::to_string(p.shape[y]));
          } else {
            conds.push_back(ix + " >= 0");
            conds.push_back(ix + " < " + std::to_string(p.shape[y]));
          }
        }
        std::string term = stride[y] == 1 ? ix : ix + " * " + std::to_string(stride[y]);
        off += (off.empty() ? "" : " + ") + term;
      }
      std::string cond;
      for (size_t c = 0; c < conds.size(); ++c) cond += (c ? " && " : "") + conds[c];
      return std::make_pair(off, cond);
    };

    if (broadcast || V == 1) {
      uniform_.insert(id);
      s << "  float " << name << "[" << U << "];\n";
      s << "#pragma unroll\n  for (int u = 0; u < " << U << "; ++u) {\n";
      auto oc = emit_rows("");
      std::string cond = oc.second.empty() ? "live[u]" : "live[u] && " + oc.second;
      s << "      " << name << "[u] = (" << cond << ") ? " << ptr << "[" << oc.first << "] : " << pad << ";\n";
      s << "  }\n";
      return;
    }
    // A row read along the innermost dimension, possibly shifted by a constant k (a translate):
    // element 0 of a vector sits at a fixed phase (k mod 4) inside a 16-byte chunk of the
    // source row, so the V elements come from V/4 (+1) aligned chunks, each entirely inside or
    // outside the row (the row length is a multiple of 4). Out-of-row chunks load from the
    // leaf's base instead and are replaced by the padding: no per-element branches. Gathers of
    // one leaf that differ only in k share their chunks (a value and its translate read the
    // same row).
    const bool window = contiguous && V % 4 == 0 && p.shape[moving] % 4 == 0 && rows[moving].others_by4 &&
                        !rows[moving].upper_only;
    std::vector<std::string> chunk(static_cast<size_t>(V / 4 + 1));
    std::vector<bool> chunk_new(chunk.size(), false);
    int phase = 0, nch = 0;
    if (window) {
      const Row &mv = rows[moving];
      phase = static_cast<int>(((mv.shift % 4) + 4) % 4);
      nch = V / 4 + (phase ? 1 : 0);
      std::string key = ptr;
      for (int y = 0; y < src_rank; ++y) key += "|" + (rows[y].empty ? std::string("_") : y == moving ? rows[y].unshifted : rows[y].expr);
      for (int q = 0; q < nch; ++q) {
        // chunk q starts at 4q - phase from element 0, i.e. at 4q - phase + k from the unshifted index
        const std::string ck = key + "|" + std::to_string(4 * q - phase + mv.shift);
        auto it = chunks_.find(ck);
        if (it != chunks_.end()) {
          chunk[q] = it->second;
        } else {
          chunk[q] = g + "_w" + std::to_string(q);
          chunk_new[q] = true;
          chunks_[ck] = chunk[q];
          s << "  csmi_f4 " << chunk[q] << "[" << U << "];\n";
        }
      }
    }
    s << "  float " << name << "[" << U << "][" << V << "];\n";
    s << "#pragma unroll\n  for (int u = 0; u < " << U << "; ++u) {\n";
    if (contiguous) {
      // One wide load when all V elements are inside the source row, else per element.
      auto oc = emit_rows("");
      std::string yi = g + "_i" + std::to_string(moving);
      std::vector<std::string> conds{"live[u]"};
      for (int y = 0; y < src_rank; ++y) {
        if (rows[y].empty || y == moving) continue;
        std::string ix = g + "_i" + std::to_string(y);
        if (rows[y].upper_only) {
          conds.push_back(ix + " < " + std::to_string(p.shape[y]));
        } else {
          conds.push_back(ix + " >= 0");
          conds.push_back(ix + " < " + std::to_string(p.shape[y]));
        }
      }
      std::string others;
      for (size_t c = 0; c < conds.size(); ++c) others += (c ? " && " : "") + conds[c];
      s << "      const bool " << g << "_ok = " << others << ";\n";
      s << "      const " << gq_ << "float* " << g << "_p = " << ptr << " + (" << oc.first << ");\n";
      if (window) {
        for (int q = 0; q < nch; ++q) {
          if (!chunk_new[q]) continue;
          const std::string cs = g + "_s" + std::to_string(q), ck = g + "_k" + std::to_string(q);
          const std::string w = chunk[q];
          const int off = 4 * q - phase;
          s << "      const int " << cs << " = " << yi << " + " << off << ";\n";
          s << "      const bool " << ck << " = " << g << "_ok && " << cs << " >= 0 && " << cs << " < "
            << p.shape[moving] << ";\n";
          s << "      const csmi_f4 " << w << "l = *(const " << gq_ << "csmi_f4*)(" << ck << " ? " << g << "_p + "
            << off << " : " << ptr << ");\n";
          s << "      " << w << "[u] = " << ck << " ? " << w << "l : csmi_f4{" << pad << ", " << pad << ", " << pad
            << ", " << pad << "};\n";
        }
        for (int v = 0; v < V; ++v)
          s << "      " << name << "[u][" << v << "] = " << chunk[(phase + v) / 4] << "[u][" << (phase + v) % 4 << "];\n";
        s << "  }\n";
        return;
      }
      s << "      if (" << g << "_ok && " << yi << " >= 0 && " << yi << " + " << (V - 1) << " < " << p.shape[moving]
        << ") {\n";
      for (int h = 0; h < V; h += 4) {
        if (V - h >= 4) {
          s << "        const csmi_f4 a" << h << "_ = *(const " << gq_ << "csmi_f4*)(" << g << "_p + " << h << ");\n";
          for (int v = 0; v < 4; ++v)
            s << "        " << name << "[u][" << h + v << "] = a" << h << "_[" << v << "];\n";
        } else {
          s << "        const csmi_f2 a" << h << "_ = *(const " << gq_ << "csmi_f2*)(" << g << "_p + " << h << ");\n";
          for (int v = 0; v < 2; ++v)
            s << "        " << name << "[u][" << h + v << "] = a" << h << "_[" << v << "];\n";
        }
      }
      s << "      } else {\n#pragma unroll\n        for (int v = 0; v < " << V << "; ++v) {\n";
      s << "          const int j_ = " << yi << " + v;\n";
      s << "          " << name << "[u][v] = (" << g << "_ok && j_ >= 0 && j_ < " << p.shape[moving] << ") ? " << g
        << "_p[v] : " << pad << ";\n";
      s << "        }\n      }\n  }\n";
      return;
    }
    // General gather: every element computes its own index.
    s << "#pragma unroll\n    for (int v = 0; v < " << V << "; ++v) {\n";
    std::string cin = "(c" + std::to_string(r - 1) + "[u] + v)";
    auto oc = emit_rows(cin);
    std::string cond = oc.second.empty() ? "live[u]" : "live[u] && " + oc.second;
    s << "      " << name << "[u][v] = (" << cond << ") ? " << ptr << "[" << oc.first << "] : " << pad << ";\n";
    s << "    }\n  }\n";
  }

  static std::string replace_coord(const std::string &e, const std::string &from, const std::string &to) {
    std::string out;
    size_t pos = 0;
    while (pos < e.size()) {
      if (e.compare(pos, from.size(), from) == 0) {
        out += to;
        pos += from.size();
      } else {
        out += e[pos++];
      }
    }
    return out;
  }

  const Dag &dag_;
  Options o_;
  std::string gq_;                    // pointer qualifier: "" in kernels, global in batch functions
  std::string bdim_ = "blockDim.x";  // workgroup size expression
  std::ostringstream *body_ = nullptr;
  std::set<int> uniform_;
  std::map<int, int> fill_of_;
  std::map<std::string, std::string> chunks_;  // aligned source chunks already loaded, by position
  std::map<std::pair<int, int64_t>, int64_t> staged_;  // (leaf, offset) -> its slot in csmi_k
  std::string decls_;                                   // module-scope declarations the body needs
  std::string prefix_;                                  // kernel name, for the body's module-scope names
  std::map<std::string, RowGroup> row_groups_;          // rows read together, by leaf and leading indices
  std::map<int, std::string> in_group_;                 // gather node -> its row group
};

// FAST mirrors -cl-unsafe-math-optimizations (OpenCL.scala:1131-1135): reassociation,
// reciprocals, no signed zeros and contraction, but not finite-math-only, so infinities and
// NaN keep their meaning (fmaxf/fminf still return the non-NaN operand).
std::vector<std::string> math_options(int mode) {
  if (mode == CS_MATH_FAST) return {"-funsafe-math-optimizations", "-ffp-contract=fast"};
  return {"-ffp-contract=off", "-fno-fast-math"};
}

}  // namespace
}  // namespace csmi

using namespace csmi;

// A loaded module that several compiled expressions may share (cs_fused_compile_many).
struct ModuleRef {
  hipModule_t module = nullptr;
  int ordinal = 0;
  ~ModuleRef() {
    if (module) {
      (void)hipSetDevice(ordinal);
      (void)hipModuleUnload(module);
    }
  }
};

struct cs_fused {
  std::atomic<int> refs{1};
  cs_device *device = nullptr;
  std::string key;
  uint64_t hash = 0;
  std::string source;
  std::shared_ptr<ModuleRef> module;
  hipFunction_t function = nullptr;
  int nleaves = 0;
  int64_t out_elements = 0;  // coordinates x tuple width
  int64_t vectors = 0;       // V-wide vectors the launch covers (the output, padded when tiled)
  uint32_t block = 256;
  uint32_t grid = 0;
  int max_block = 1024;      // the kernel's launch bound
  Dag dag;                   // kept for batches, which re-emit the body
  Options opts;
};

struct cs_batch {
  cs_device *device = nullptr;
  std::vector<cs_fused *> exprs;
  std::string source;
  hipModule_t module = nullptr;
  hipFunction_t kernel = nullptr, reset = nullptr;
  uint32_t block = 256;
  uint32_t grid = 0;
  int64_t items = 0;
  int max_args = 0;
  int total_args = 0;
  void *items_dev = nullptr;  // uint2 {expr, tile} per work item
  void *args_dev = nullptr;   // exprs x max_args device pointers
  unsigned *counter = nullptr;
};

namespace {

// Kernel symbol of a compiled expression: its structural hash and launch shape, so that
// profiles tell expressions and configurations apart.
std::string kernel_name(const std::string &key, const Options &o) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "csmi_fused_%016llx_v%du%d%s", static_cast<unsigned long long>(fnv1a(key)), o.V, o.U,
                o.sum ? "_sum" : "");
  return buf;
}

Dag prepare(const cs_node *nodes, int nnodes, const int32_t *outputs, int noutputs, int out_rank,
            const int32_t *out_shape, const cs_fused_options *options, Options *o, std::string *key) {
  Dag dag = canonicalize(nodes, nnodes, outputs, noutputs, out_rank, out_shape);
  *o = resolve_options(options, dag);
  *key = canonical_key(dag, *o);
  return dag;
}

// One member of a batch: its canonical DAG and resolved launch options.
struct BatchExpr {
  const Dag *dag;
  Options opts;
};

// The persistent batch kernel's source and its work items: (expression, first tile, tile
// count, 0) per item, expressions with the most work per tile first.
struct BatchPlan {
  std::string source;
  std::vector<uint32_t> items;
  int max_args = 1;
};

BatchPlan plan_batch(const std::vector<BatchExpr> &members, int block, int math_mode) {
  const int n = static_cast<int>(members.size());
  BatchPlan plan;
  std::ostringstream src;
  src << "// csmi batch of " << n << " expressions\n" << Emitter::prelude(math_mode);
  std::vector<int64_t> tiles(static_cast<size_t>(n)), tile_bytes(static_cast<size_t>(n));
  std::vector<double> cost(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    const Dag &dag = *members[i].dag;
    const Options &o = members[i].opts;
    int64_t vectors = 0;
    src << Emitter(dag, o).device_function("expr_" + std::to_string(i), block, &vectors);
    int64_t per = int64_t(block) * o.U;
    tiles[i] = (vectors + per - 1) / per;
    tile_bytes[i] = per * o.V * int64_t(dag.outputs.size() + dag.nleaves) * 4;
    cost[i] = double(vectors) * o.V * (dag.nleaves + 1 + double(dag.nodes.size()) / 4.0);
    plan.max_args = std::max(plan.max_args, dag.nleaves + 1);
  }
  src << "extern \"C\" __global__ void batch_reset(unsigned* counter) { *counter = 0u; }\n";
  // Work item = (expression, first tile, tile count). Thread 0 takes the next item from
  // the counter while the block works on the current one, hiding the atomic's latency.
  src << "extern \"C\" __global__ __launch_bounds__(" << block << ") void batch_kernel(const uint4* __restrict__ items, "
         "const unsigned nitems, const float* const* __restrict__ args, unsigned* counter) {\n";
  src << "  __shared__ unsigned next;\n";
  src << "  if (threadIdx.x == 0) next = atomicAdd(counter, 1u);\n  __syncthreads();\n";
  // The item index is the same in every lane: readfirstlane makes it scalar, so the item
  // and its argument pointers come through scalar loads.
  src << "  unsigned w = __builtin_amdgcn_readfirstlane(next);\n  __syncthreads();\n";
  src << "  while (w < nitems) {\n";
  src << "    if (threadIdx.x == 0) next = atomicAdd(counter, 1u);\n";
  src << "    const uint4 it = items[w];\n    const float* const* a = args + it.x * " << plan.max_args << "u;\n";
  src << "    switch (it.x) {\n";
  for (int i = 0; i < n; ++i) {
    const int nl = members[i].dag->nleaves;
    src << "      case " << i << ": expr_" << i << "(it.y, it.y + it.z";
    for (int k = 0; k < nl; ++k) src << ", (const CSMI_G float*)a[" << k << "]";
    src << ", (CSMI_G float*)a[" << nl << "]); break;\n";
  }
  src << "      default: break;\n    }\n";
  src << "    __syncthreads();\n    w = __builtin_amdgcn_readfirstlane(next);\n    __syncthreads();\n  }\n}\n";
  plan.source = src.str();
  std::vector<int> order(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return cost[x] / std::max<int64_t>(tiles[x], 1) > cost[y] / std::max<int64_t>(tiles[y], 1);
  });
  // Runs of consecutive tiles of about kItemBytes of traffic: few enough items that the
  // single counter is not a bottleneck, many enough to balance the grid.
  constexpr int64_t kItemBytes = 128 << 10;
  for (int i : order) {
    int64_t run = std::max<int64_t>(1, kItemBytes / std::max<int64_t>(tile_bytes[i], 1));
    for (int64_t t = 0; t < tiles[i]; t += run) {
      plan.items.push_back(static_cast<uint32_t>(i));
      plan.items.push_back(static_cast<uint32_t>(t));
      plan.items.push_back(static_cast<uint32_t>(std::min(run, tiles[i] - t)));
      plan.items.push_back(0u);
    }
  }
  return plan;
}

}  // namespace

static cs_status compile_fused(cs_device *device, const cs_node *nodes, int nnodes, const int32_t *outputs,
                               int noutputs, int out_rank, const int32_t *out_shape,
                               const cs_fused_options *options, bool sum, cs_fused **fused) {
  return guard([&] {
    if (!device) fail(CS_INVALID_DEVICE, "device is null");
    if (!fused) fail(CS_INVALID_VALUE, "fused is null");
    Options o;
    std::string key;
    Dag dag = canonicalize(nodes, nnodes, outputs, noutputs, out_rank, out_shape);
    o = resolve_options(options, dag);
    o.sum = sum;
    if (sum && dag.outputs.size() != 1) fail(CS_INVALID_VALUE, "a fused sum takes one output");
    key = canonical_key(dag, o);
    {
      std::lock_guard<std::recursive_mutex> l(device->mu);
      auto it = device->fused_cache.find(key);
      if (it != device->fused_cache.end()) {
        it->second->refs++;
        *fused = it->second;
        return;
      }
    }
    int64_t vectors = 0;
    const std::string name = kernel_name(key, o);
    std::string src = Emitter(dag, o).generate(&vectors, name);
    std::string log;
    auto mod = std::make_shared<ModuleRef>();
    mod->ordinal = device->ordinal;
    mod->module = build_module(device, src, math_options(o.math_mode), &log);
    std::unique_ptr<cs_fused> f(new cs_fused());
    f->device = device;
    f->key = key;
    f->hash = fnv1a(key);
    f->source = src;
    f->module = mod;
    CS_HIP(hipModuleGetFunction(&f->function, mod->module, name.c_str()));
    f->nleaves = dag.nleaves;
    int64_t total = 1;
    for (auto d : dag.out_shape) total *= d;
    f->out_elements = total * static_cast<int64_t>(dag.outputs.size());
    f->vectors = vectors;
    f->block = static_cast<uint32_t>(o.block);
    f->max_block = launch_bound(dag, o);
    int64_t per_block = int64_t(o.block) * o.U;
    f->grid = static_cast<uint32_t>((vectors + per_block - 1) / per_block);
    f->dag = std::move(dag);
    f->opts = o;
    std::lock_guard<std::recursive_mutex> l(device->mu);
    auto it = device->fused_cache.find(key);
    if (it != device->fused_cache.end()) {  // another thread won the race
      it->second->refs++;
      *fused = it->second;
      return;
    }
    f->refs = 2;  // one for the cache, one for the caller
    device->fused_cache.emplace(key, f.get());
    *fused = f.release();
  });
}

extern "C" {

cs_status cs_fused_generate_source(const cs_node *nodes, int nnodes, const int32_t *outputs,
                                   int noutputs, int out_rank, const int32_t *out_shape,
                                   const cs_fused_options *options, char **source, uint64_t *key) {
  return guard([&] {
    if (!source) fail(CS_INVALID_VALUE, "source is null");
    Options o;
    std::string k;
    Dag dag = prepare(nodes, nnodes, outputs, noutputs, out_rank, out_shape, options, &o, &k);
    int64_t vectors = 0;
    std::string src = Emitter(dag, o).generate(&vectors, kernel_name(k, o));
    *source = static_cast<char *>(std::malloc(src.size() + 1));
    if (!*source) fail(CS_OUT_OF_HOST_MEMORY, "malloc");
    std::memcpy(*source, src.c_str(), src.size() + 1);
    if (key) *key = fnv1a(k);
  });
}

cs_status cs_batch_generate_source(const cs_fused_desc *descs, int n, int block, char **source, int64_t *items) {
  return guard([&] {
    if (!source) fail(CS_INVALID_VALUE, "source is null");
    if (n <= 0 || !descs) fail(CS_INVALID_VALUE, "empty batch");
    if (block < 64 || block > 1024 || block % 64) fail(CS_INVALID_WORK_GROUP_SIZE, "block must be 64..1024");
    std::vector<Dag> dags(static_cast<size_t>(n));
    std::vector<BatchExpr> members;
    for (int i = 0; i < n; ++i) {
      const cs_fused_desc &d = descs[i];
      Options o;
      std::string k;
      dags[i] = prepare(d.nodes, d.nnodes, d.outputs, d.noutputs, d.out_rank, d.out_shape, &d.options, &o, &k);
      members.push_back({nullptr, o});
    }
    for (int i = 0; i < n; ++i) {
      members[i].dag = &dags[i];
      if (members[i].opts.math_mode != members[0].opts.math_mode)
        fail(CS_INVALID_VALUE, "mixed math modes in one batch");
    }
    BatchPlan plan = plan_batch(members, block, members[0].opts.math_mode);
    *source = static_cast<char *>(std::malloc(plan.source.size() + 1));
    if (!*source) fail(CS_OUT_OF_HOST_MEMORY, "malloc");
    std::memcpy(*source, plan.source.c_str(), plan.source.size() + 1);
    if (items) *items = static_cast<int64_t>(plan.items.size() / 4);
  });
}

cs_status cs_fused_precompile(const cs_node *nodes, int nnodes, const int32_t *outputs, int noutputs,
                              int out_rank, const int32_t *out_shape, const cs_fused_options *options) {
  return guard([&] {
    Options o;
    std::string k;
    Dag dag = prepare(nodes, nnodes, outputs, noutputs, out_rank, out_shape, options, &o, &k);
    int64_t vectors = 0;
    std::string src = Emitter(dag, o).generate(&vectors, kernel_name(k, o));
    std::string log;
    if (!precompile(src, math_options(o.math_mode), "gfx950", &log))
      fail(CS_BUILD_PROGRAM_FAILURE, "hiprtc build failed:\n" + log);
  });
}

cs_status cs_program_check(const char *source, const char *options, char **log) {
  if (log) *log = nullptr;
  std::string l;
  cs_status s = guard([&] {
    if (!source) fail(CS_INVALID_VALUE, "source is null");
    std::vector<std::string> opts;
    if (options) {
      std::istringstream in(options);
      std::string w;
      while (in >> w) opts.push_back(w);
    }
    size_t size = 0;
    if (!compile_check(source, opts, "gfx950", &l, &size))
      fail(CS_BUILD_PROGRAM_FAILURE, "hiprtc build failed:\n" + l);
  });
  if (log && !l.empty()) {
    *log = static_cast<char *>(std::malloc(l.size() + 1));
    if (*log) std::memcpy(*log, l.c_str(), l.size() + 1);
  }
  return s;
}

cs_status cs_fused_compile(cs_device *device, const cs_node *nodes, int nnodes, const int32_t *outputs,
                           int noutputs, int out_rank, const int32_t *out_shape, const cs_fused_options *options,
                           cs_fused **fused) {
  return compile_fused(device, nodes, nnodes, outputs, noutputs, out_rank, out_shape, options, false, fused);
}

cs_status cs_fused_compile_sum(cs_device *device, const cs_node *nodes, int nnodes, const int32_t *outputs,
                               int noutputs, int out_rank, const int32_t *out_shape,
                               const cs_fused_options *options, cs_fused **fused) {
  return compile_fused(device, nodes, nnodes, outputs, noutputs, out_rank, out_shape, options, true, fused);
}

cs_status cs_fused_sum(cs_stream *stream, cs_fused *fused, const void *const *leaves, int nleaves, float *out,
                       double *out64) {
  return guard([&] {
    if (!stream || !fused) fail(CS_INVALID_VALUE, "null stream or kernel");
    if (!fused->opts.sum) fail(CS_INVALID_KERNEL_ARGS, "the kernel was not compiled by cs_fused_compile_sum");
    if (!out && !out64) fail(CS_INVALID_VALUE, "no output");
    if (nleaves != fused->nleaves)
      fail(CS_INVALID_KERNEL_ARGS, "kernel takes " + std::to_string(fused->nleaves) + " leaves, got " +
                                       std::to_string(nleaves));
    set_device(stream->device);
    double *partials = sum_partials(stream, std::max<size_t>(fused->grid, 1));
    if (fused->grid > 0) {
      std::vector<const void *> vals(static_cast<size_t>(nleaves) + 1);
      std::vector<void *> args(static_cast<size_t>(nleaves) + 1);
      for (int i = 0; i < nleaves; ++i) {
        if (!leaves || !leaves[i]) fail(CS_INVALID_KERNEL_ARGS, "leaf " + std::to_string(i) + " is null");
        vals[i] = leaves[i];
        args[i] = &vals[i];
      }
      vals[nleaves] = partials;
      args[nleaves] = &vals[nleaves];
      CS_HIP(hipModuleLaunchKernel(fused->function, fused->grid, 1, 1, fused->block, 1, 1, 0, stream->stream,
                                   args.data(), nullptr));
    }
    launch_finish_sum(stream, partials, fused->grid, out, out64);  // zero partials: an empty sum, 0
    CS_HIP(hipGetLastError());
  });
}

// Takes one reference per entry of out[0..n); on failure drops those taken so far and leaves
// every entry null.
static void compile_many(cs_device *device, const cs_fused_desc *descs, int n, cs_fused **out);

cs_status cs_fused_compile_many(cs_device *device, const cs_fused_desc *descs, int n, cs_fused **out) {
  return guard([&] {
    if (!device) fail(CS_INVALID_DEVICE, "device is null");
    if (n < 0 || (n > 0 && (!descs || !out))) fail(CS_INVALID_VALUE, "null argument");
    for (int i = 0; i < n; ++i) out[i] = nullptr;
    try {
      compile_many(device, descs, n, out);
    } catch (...) {
      std::lock_guard<std::recursive_mutex> l(device->mu);
      for (int i = 0; i < n; ++i) {
        if (out[i] && --out[i]->refs == 0) delete out[i];
        out[i] = nullptr;
      }
      throw;
    }
  });
}

static void compile_many(cs_device *device, const cs_fused_desc *descs, int n, cs_fused **out) {
  {
    std::vector<Dag> dags(static_cast<size_t>(n));
    std::vector<Options> opts(static_cast<size_t>(n));
    std::vector<std::string> keys(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
      const cs_fused_desc &d = descs[i];
      dags[i] = prepare(d.nodes, d.nnodes, d.outputs, d.noutputs, d.out_rank, d.out_shape, &d.options, &opts[i],
                        &keys[i]);
    }
    // Cached expressions come from the cache; the rest compile together, one module per
    // math mode (the modes differ in compiler options).
    std::map<int, std::vector<int>> todo;
    {
      std::lock_guard<std::recursive_mutex> l(device->mu);
      for (int i = 0; i < n; ++i) {
        auto it = device->fused_cache.find(keys[i]);
        if (it != device->fused_cache.end()) {
          it->second->refs++;
          out[i] = it->second;
        } else {
          bool dup = false;
          for (int j : todo[opts[i].math_mode]) dup = dup || keys[j] == keys[i];
          if (!dup) todo[opts[i].math_mode].push_back(i);
        }
      }
    }
    for (auto &group : todo) {
      std::ostringstream src;
      src << Emitter::prelude(group.first);
      std::vector<int64_t> vectors(group.second.size());
      for (size_t k = 0; k < group.second.size(); ++k) {
        int i = group.second[k];
        src << Emitter(dags[i], opts[i]).generate(&vectors[k], kernel_name(keys[i], opts[i]), false);
      }
      std::string log;
      auto mod = std::make_shared<ModuleRef>();
      mod->ordinal = device->ordinal;
      mod->module = build_module(device, src.str(), math_options(group.first), &log);
      for (size_t k = 0; k < group.second.size(); ++k) {
        int i = group.second[k];
        std::unique_ptr<cs_fused> f(new cs_fused());
        f->device = device;
        f->key = keys[i];
        f->hash = fnv1a(keys[i]);
        int64_t v = 0;
        const std::string name = kernel_name(keys[i], opts[i]);
        f->source = Emitter(dags[i], opts[i]).generate(&v, name);
        f->module = mod;
        CS_HIP(hipModuleGetFunction(&f->function, mod->module, name.c_str()));
        f->nleaves = dags[i].nleaves;
        int64_t total = 1;
        for (auto d : dags[i].out_shape) total *= d;
        f->out_elements = total * static_cast<int64_t>(dags[i].outputs.size());
        f->vectors = vectors[k];
        f->block = static_cast<uint32_t>(opts[i].block);
        f->max_block = launch_bound(dags[i], opts[i]);
        int64_t per_block = int64_t(opts[i].block) * opts[i].U;
        f->grid = static_cast<uint32_t>((vectors[k] + per_block - 1) / per_block);
        f->dag = dags[i];
        f->opts = opts[i];
        std::lock_guard<std::recursive_mutex> l(device->mu);
        auto it = device->fused_cache.find(keys[i]);
        if (it != device->fused_cache.end()) {
          it->second->refs++;
          out[i] = it->second;
        } else {
          f->refs = 2;
          device->fused_cache.emplace(keys[i], f.get());
          out[i] = f.release();
        }
      }
    }
    // duplicates within the request resolve to the now-cached kernel
    std::lock_guard<std::recursive_mutex> l(device->mu);
    for (int i = 0; i < n; ++i)
      if (!out[i]) {
        auto it = device->fused_cache.find(keys[i]);
        if (it == device->fused_cache.end()) fail(CS_INVALID_OPERATION, "expression missing after compile");
        it->second->refs++;
        out[i] = it->second;
      }
  }
}

cs_status cs_fused_launch_block(cs_stream *stream, cs_fused *fused, const void *const *leaves, int nleaves,
                                void *out, int block) {
  return guard([&] {
    if (!stream || !fused) fail(CS_INVALID_VALUE, "null stream or kernel");
    if (fused->opts.sum) fail(CS_INVALID_KERNEL_ARGS, "a sum kernel runs through cs_fused_sum");
    if (block == 0) block = static_cast<int>(fused->block);
    if (block < 64 || block > 1024 || block % 64)
      fail(CS_INVALID_WORK_GROUP_SIZE, "block must be a multiple of 64 in 64..1024");
    if (block > fused->max_block)
      fail(CS_INVALID_WORK_GROUP_SIZE, "this expression's kernel was compiled for at most " +
                                           std::to_string(fused->max_block) + " threads per workgroup");
    if (nleaves != fused->nleaves)
      fail(CS_INVALID_KERNEL_ARGS, "kernel takes " + std::to_string(fused->nleaves) + " leaves, got " +
                                       std::to_string(nleaves));
    int64_t per_block = int64_t(block) * fused->opts.U;
    uint32_t grid = static_cast<uint32_t>((fused->vectors + per_block - 1) / per_block);
    if (grid == 0) return;
    if (!out) fail(CS_INVALID_KERNEL_ARGS, "output pointer is null");
    std::vector<const void *> vals(static_cast<size_t>(nleaves) + 1);
    std::vector<void *> args(static_cast<size_t>(nleaves) + 1);
    for (int i = 0; i < nleaves; ++i) {
      if (!leaves || !leaves[i]) fail(CS_INVALID_KERNEL_ARGS, "leaf " + std::to_string(i) + " is null");
      vals[i] = leaves[i];
      args[i] = &vals[i];
    }
    vals[nleaves] = out;
    args[nleaves] = &vals[nleaves];
    set_device(stream->device);
    CS_HIP(hipModuleLaunchKernel(fused->function, grid, 1, 1, static_cast<uint32_t>(block), 1, 1, 0,
                                 stream->stream, args.data(), nullptr));
  });
}

cs_status cs_fused_retain(cs_fused *fused) {
  return guard([&] {
    if (!fused) fail(CS_INVALID_VALUE, "fused is null");
    fused->refs++;
  });
}

cs_status cs_fused_release(cs_fused *fused) {
  return guard([&] {
    if (!fused) fail(CS_INVALID_VALUE, "fused is null");
    if (--fused->refs == 0) delete fused;  // the module goes with its last expression
  });
}

cs_status cs_fused_source(cs_fused *fused, const char **source) {
  return guard([&] {
    if (!fused || !source) fail(CS_INVALID_VALUE, "null argument");
    *source = fused->source.c_str();
  });
}

cs_status cs_fused_signature(cs_fused *fused, int *nleaves, int64_t *out_elements) {
  return guard([&] {
    if (!fused) fail(CS_INVALID_VALUE, "fused is null");
    if (nleaves) *nleaves = fused->nleaves;
    if (out_elements) *out_elements = fused->out_elements;
  });
}

cs_status cs_fused_key(cs_fused *fused, uint64_t *key) {
  return guard([&] {
    if (!fused || !key) fail(CS_INVALID_VALUE, "null argument");
    *key = fused->hash;
  });
}

cs_status cs_fused_launch(cs_stream *stream, cs_fused *fused, const void *const *leaves, int nleaves,
                          void *out) {
  return guard([&] {
    if (!stream || !fused) fail(CS_INVALID_VALUE, "null stream or kernel");
    if (fused->opts.sum) fail(CS_INVALID_KERNEL_ARGS, "a sum kernel runs through cs_fused_sum");
    if (nleaves != fused->nleaves)
      fail(CS_INVALID_KERNEL_ARGS, "kernel takes " + std::to_string(fused->nleaves) + " leaves, got " +
                                       std::to_string(nleaves));
    if (fused->grid == 0) return;
    if (!out) fail(CS_INVALID_KERNEL_ARGS, "output pointer is null");
    std::vector<const void *> vals(static_cast<size_t>(nleaves) + 1);
    std::vector<void *> args(static_cast<size_t>(nleaves) + 1);
    for (int i = 0; i < nleaves; ++i) {
      if (!leaves || !leaves[i]) fail(CS_INVALID_KERNEL_ARGS, "leaf " + std::to_string(i) + " is null");
      vals[i] = leaves[i];
      args[i] = &vals[i];
    }
    vals[nleaves] = out;
    args[nleaves] = &vals[nleaves];
    set_device(stream->device);
    CS_HIP(hipModuleLaunchKernel(fused->function, fused->grid, 1, 1, fused->block, 1, 1, 0,
                                 stream->stream, args.data(), nullptr));
  });
}

cs_status cs_fused_launch_timed(cs_stream *stream, cs_fused *fused, const void *const *leaves, int nleaves,
                                void *out, cs_event **start, cs_event **stop) {
  return guard([&] {
    if (!stream || !fused) fail(CS_INVALID_VALUE, "null stream or kernel");
    if (fused->opts.sum) fail(CS_INVALID_KERNEL_ARGS, "a sum kernel runs through cs_fused_sum");
    if (!start || !stop) fail(CS_INVALID_VALUE, "null event pointer");
    if (nleaves != fused->nleaves)
      fail(CS_INVALID_KERNEL_ARGS, "kernel takes " + std::to_string(fused->nleaves) + " leaves, got " +
                                       std::to_string(nleaves));
    if (fused->grid == 0) fail(CS_INVALID_OPERATION, "an empty output launches no kernel to time");
    if (!out) fail(CS_INVALID_KERNEL_ARGS, "output pointer is null");
    std::vector<const void *> vals(static_cast<size_t>(nleaves) + 1);
    std::vector<void *> args(static_cast<size_t>(nleaves) + 1);
    for (int i = 0; i < nleaves; ++i) {
      if (!leaves || !leaves[i]) fail(CS_INVALID_KERNEL_ARGS, "leaf " + std::to_string(i) + " is null");
      vals[i] = leaves[i];
      args[i] = &vals[i];
    }
    vals[nleaves] = out;
    args[nleaves] = &vals[nleaves];
    set_device(stream->device);
    hipEvent_t ea = nullptr, eb = nullptr;
    hipError_t r = hipEventCreate(&ea);
    if (r == hipSuccess) r = hipEventCreate(&eb);
    // hipExtModuleLaunchKernel takes the grid in work items; the events carry the
    // dispatch's own start and end
    if (r == hipSuccess)
      r = hipExtModuleLaunchKernel(fused->function, fused->grid * fused->block, 1, 1, fused->block, 1, 1, 0,
                                   stream->stream, args.data(), nullptr, ea, eb, 0);
    if (r != hipSuccess) {
      if (ea) (void)hipEventDestroy(ea);
      if (eb) (void)hipEventDestroy(eb);
      CS_HIP(r);
    }
    auto *a = new cs_event(), *b = new cs_event();
    a->ordinal = b->ordinal = stream->device->ordinal;
    a->event = ea;
    b->event = eb;
    *start = a;
    *stop = b;
  });
}

cs_status cs_fused_lookup(cs_device *device, const cs_node *nodes, int nnodes, const int32_t *outputs,
                          int noutputs, int out_rank, const int32_t *out_shape,
                          const cs_fused_options *options, int *hit) {
  return guard([&] {
    if (!device || !hit) fail(CS_INVALID_VALUE, "null argument");
    Options o;
    std::string key;
    prepare(nodes, nnodes, outputs, noutputs, out_rank, out_shape, options, &o, &key);
    std::lock_guard<std::recursive_mutex> l(device->mu);
    *hit = device->fused_cache.count(key) ? 1 : 0;
  });
}

cs_status cs_fused_cache_clear(cs_device *device) {
  return guard([&] {
    if (!device) fail(CS_INVALID_DEVICE, "device is null");
    std::vector<cs_fused *> drop;
    {
      std::lock_guard<std::recursive_mutex> l(device->mu);
      for (auto &kv : device->fused_cache) drop.push_back(kv.second);
      device->fused_cache.clear();
    }
    for (cs_fused *f : drop) cs_fused_release(f);
  });
}

// ------------------------------------------------------------------------ batches
// One persistent kernel for many fused expressions: every expression's body becomes a
// device function over a tile index (its own vector width and vectors per thread), and a
// queue of (expression, tile) work items, ordered by decreasing expression cost, is drained
// by a fixed grid of workgroups through an atomic counter. A reset kernel zeroes the counter
// before each run, so a captured graph replays correctly.
cs_status cs_batch_create(cs_device *device, cs_fused *const *exprs, int n, int block, cs_batch **batch) {
  return guard([&] {
    if (!device || !batch) fail(CS_INVALID_VALUE, "null argument");
    if (n <= 0 || !exprs) fail(CS_INVALID_VALUE, "empty batch");
    if (block < 64 || block > 1024 || block % 64) fail(CS_INVALID_WORK_GROUP_SIZE, "block must be 64..1024");
    for (int i = 0; i < n; ++i)
      if (!exprs[i]) fail(CS_INVALID_VALUE, "null expression");
    int math_mode = exprs[0]->opts.math_mode;
    for (int i = 0; i < n; ++i) {
      if (exprs[i]->device != device) fail(CS_INVALID_DEVICE, "expression compiled on another device");
      if (exprs[i]->opts.math_mode != math_mode) fail(CS_INVALID_VALUE, "mixed math modes in one batch");
      if (exprs[i]->opts.sum) fail(CS_INVALID_KERNEL_ARGS, "a sum kernel cannot join a batch");
    }
    std::vector<BatchExpr> members;
    for (int i = 0; i < n; ++i) members.push_back({&exprs[i]->dag, exprs[i]->opts});
    BatchPlan plan = plan_batch(members, block, math_mode);
    std::unique_ptr<cs_batch, void (*)(cs_batch *)> b(new cs_batch(), [](cs_batch *x) {
      if (x->module) (void)hipModuleUnload(x->module);
      (void)hipFree(x->items_dev);
      (void)hipFree(x->args_dev);
      (void)hipFree(x->counter);
      delete x;
    });
    b->device = device;
    b->block = static_cast<uint32_t>(block);
    b->source = plan.source;
    b->max_args = plan.max_args;
    std::string log;
    set_device(device);
    b->module = build_module(device, b->source, math_options(math_mode), &log);
    CS_HIP(hipModuleGetFunction(&b->kernel, b->module, "batch_kernel"));
    CS_HIP(hipModuleGetFunction(&b->reset, b->module, "batch_reset"));
    const std::vector<uint32_t> &items = plan.items;
    b->items = static_cast<int64_t>(items.size() / 4);
    CS_HIP(hipMalloc(&b->items_dev, std::max<size_t>(items.size() * 4, 8)));
    if (!items.empty()) CS_HIP(hipMemcpy(b->items_dev, items.data(), items.size() * 4, hipMemcpyHostToDevice));
    const size_t table = size_t(n) * size_t(b->max_args) * sizeof(void *);
    CS_HIP(hipMalloc(&b->args_dev, table));
    CS_HIP(hipMemset(b->args_dev, 0, table));
    CS_HIP(hipMalloc(reinterpret_cast<void **>(&b->counter), 256));
    CS_HIP(hipMemset(b->counter, 0, 256));
    int occupancy = 0;
    CS_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occupancy, b->kernel, block, 0));
    if (occupancy < 1) occupancy = 1;
    int64_t want = int64_t(device->props.multiProcessorCount) * occupancy;
    b->grid = static_cast<uint32_t>(std::max<int64_t>(1, std::min<int64_t>(want, b->items)));
    for (int i = 0; i < n; ++i) {
      exprs[i]->refs++;
      b->exprs.push_back(exprs[i]);
    }
    for (int i = 0; i < n; ++i) b->total_args += exprs[i]->nleaves + 1;
    *batch = b.release();
  });
}

cs_status cs_batch_bind(cs_stream *stream, cs_batch *batch, const void *const *args, int nargs) {
  return guard([&] {
    if (!stream || !batch) fail(CS_INVALID_VALUE, "null argument");
    if (nargs != batch->total_args)
      fail(CS_INVALID_KERNEL_ARGS, "batch takes " + std::to_string(batch->total_args) + " pointers");
    std::vector<const void *> table(batch->exprs.size() * size_t(batch->max_args), nullptr);
    int k = 0;
    for (size_t i = 0; i < batch->exprs.size(); ++i)
      for (int j = 0; j <= batch->exprs[i]->nleaves; ++j) {
        if (!args || !args[k]) fail(CS_INVALID_KERNEL_ARGS, "null pointer in batch arguments");
        table[i * batch->max_args + j] = args[k++];
      }
    set_device(stream->device);
    CS_HIP(hipMemcpyAsync(batch->args_dev, table.data(), table.size() * sizeof(void *), hipMemcpyHostToDevice,
                          stream->stream));
    CS_HIP(hipStreamSynchronize(stream->stream));
  });
}

cs_status cs_batch_launch(cs_stream *stream, cs_batch *batch) {
  return guard([&] {
    if (!stream || !batch) fail(CS_INVALID_VALUE, "null argument");
    set_device(stream->device);
    void *counter = batch->counter;
    void *rargs[] = {&counter};
    CS_HIP(hipModuleLaunchKernel(batch->reset, 1, 1, 1, 1, 1, 1, 0, stream->stream, rargs, nullptr));
    if (batch->items == 0) return;
    const void *items = batch->items_dev;
    unsigned nitems = static_cast<unsigned>(batch->items);
    const void *argtab = batch->args_dev;
    void *kargs[] = {&items, &nitems, &argtab, &counter};
    CS_HIP(hipModuleLaunchKernel(batch->kernel, batch->grid, 1, 1, batch->block, 1, 1, 0, stream->stream, kargs,
                                 nullptr));
  });
}

cs_status cs_batch_info(cs_batch *batch, int64_t *items, int *grid, int *block) {
  return guard([&] {
    if (!batch) fail(CS_INVALID_VALUE, "batch is null");
    if (items) *items = batch->items;
    if (grid) *grid = static_cast<int>(batch->grid);
    if (block) *block = static_cast<int>(batch->block);
  });
}

cs_status cs_batch_release(cs_batch *batch) {
  return guard([&] {
    if (!batch) fail(CS_INVALID_VALUE, "batch is null");
    set_device(batch->device);
    (void)hipModuleUnload(batch->module);
    (void)hipFree(batch->items_dev);
    (void)hipFree(batch->args_dev);
    (void)hipFree(batch->counter);
    for (cs_fused *f : batch->exprs) cs_fused_release(f);
    delete batch;
  });
}

}  // extern "C"
