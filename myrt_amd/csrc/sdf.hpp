// Signed-distance programs: the tape format shared by the host (scene.cpp) and sdf.hip.
//
// The reference compiles an sdf:: tree of primitives, operators and modifiers into a GLSL map()
// function by string generation (src/pathtracer/sdf.cpp:122-407). Here the tree is a postfix
// tape of 32-byte instructions evaluated by one device interpreter: primitives push a distance
// (and material), operators pop two and push one, PUSH_* modifiers save the position and
// transform it for the following subtree, POP restores it.
#pragma once
#include <cstdint>

namespace mrt {

enum SdfOp : uint32_t {
  kSdfSphere = 1,       // p0 radius                         (sdf.hpp:399-414)
  kSdfBox = 2,          // p0-p2 half extents                 (sdf.hpp:524-542)
  kSdfTorus = 3,        // p0 large radius, p1 small radius   (sdf.hpp:545-565)
  kSdfCapsule = 4,      // vertical: p0 height, p1 radius     (sdf.hpp:603-618)
  kSdfMenger = 5,       // unit Menger sponge, 4 iterations   (sdf.hpp:567-601)
  kSdfUnion = 16,       // min(a, b)                          (sdf.hpp:429-438)
  kSdfSubtract = 17,    // max(-a, b)                         (sdf.hpp:440-449)
  kSdfIntersect = 18,   // max(a, b)                          (sdf.hpp:451-460)
  kSdfSmoothUnion = 19,      // p0 k                          (sdf.hpp:462-482)
  kSdfSmoothSubtract = 20,   // p0 k                          (sdf.hpp:484-502)
  kSdfSmoothIntersect = 21,  // p0 k                          (sdf.hpp:504-522)
  kSdfPushTranslate = 32,    // p -= (p0, p1, p2)             (sdf.hpp:416-427)
  kSdfPushMirror = 33,       // p[p0] = |p[p0]|               (sdf.hpp:620-632)
  kSdfPop = 34,              // restore the saved position
};

struct SdfNode {
  uint32_t op;
  uint32_t material;
  float p[6];
};
static_assert(sizeof(SdfNode) == 32, "sdf node is 32 bytes");

// Executable instruction, 48 B: the host compiles each tape into these before upload. Bits 0-7 of
// `op` hold a tape op. A primitive may carry a translation (a PUSH_TRANSLATE, primitive, POP
// triple of the tape; bit kSdfFuseTranslate) and a binary operator that directly follows it on
// the tape (bits 8-15): the operator then combines the stack top (lhs) with the primitive (rhs)
// in place. Fusion keeps every binary32 operation of the tape in its order, so the results are
// those of the plain tape; it removes most stack shifts and instruction dispatches.
struct SdfInsn {
  uint32_t op;
  uint32_t material;
  float t[3];   // translation subtracted from the position before the primitive
  float k;      // parameter of the fused operator
  float p[6];   // parameters of the tape op
};
static_assert(sizeof(SdfInsn) == 48, "sdf instruction is 48 bytes");
constexpr uint32_t kSdfFuseTranslate = 1u << 16;

constexpr int kSdfStack = 8;     // distance/material stack depth of the interpreter
constexpr int kSdfPosStack = 4;  // nested modifier depth

// One enqueued SDF drawable (drawable_sdf_t, src/pt/pt_trace.comp:49-54), 80 B.
struct SdfDrawable {
  float inv[12];      // rows 0-2 of M^-1
  float box_lo[3];    // world-space bounds of the surface (for ray clipping)
  uint32_t tape_offset;
  float box_hi[3];
  uint32_t tape_len;
};
static_assert(sizeof(SdfDrawable) == 80, "sdf drawable is 80 bytes");

// Culling grid of a program that is one chain of primitives joined by (smooth) unions: for each
// cell of a regular grid over the program's object-space bounds, the set of chain instructions
// (bit j = instruction j, bit 0 = the head) whose primitive can change the distance somewhere in
// the cell. A dropped primitive is at least k (plus a rounding margin) above the running value
// everywhere in the cell, where the union returns its left operand bit for bit.
struct SdfGrid {
  float lo[3];
  uint32_t dim;          // cells per axis; 0 = no grid (every instruction is evaluated)
  float inv[3];          // cells per object-space unit
  uint32_t mask_offset;  // first cell's mask in the mask array
};
constexpr int kSdfGridCoarse = 32;  // cells per axis of the first pass of the grid build
constexpr int kSdfChainMax = 64;

}  // namespace mrt
