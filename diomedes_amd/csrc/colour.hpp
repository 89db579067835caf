// colour.hpp — graph colouring of the solve's body graph for the coloured
// Gauss-Seidel primal (no counterpart in the reference, whose solve is
// Jacobi).
//
// Vertices are the dynamic bodies; two are adjacent when they share a
// contact or a joint.  Jones-Plassmann rounds with hashed priorities: in
// round r every uncoloured body whose priority beats each uncoloured
// neighbour's takes the smallest colour that no neighbour coloured in an
// earlier round holds.  A body coloured in round r records r, and readers
// count only rounds < r as coloured, so a round's decisions depend only on
// the state before it: the colouring is the same however the threads run.
// Then the bodies are sorted by colour (stable: ascending index within a
// colour) into per-colour lists.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dio {

constexpr uint32_t kMaxColours = 255;  // colours 0 .. 254; sort key 255 = not in the graph
constexpr uint32_t kColourUncoloured = 0xFFFFFFFFu, kColourNotInGraph = 0xFFFFFFFEu;
constexpr uint32_t kColourRoundSlots = 4096;  // rounds the pending flags cover

// The graph: the step's contact CSR (with peer codes) and joint CSR.
struct ColourGraph {
    uint32_t n = 0;
    const float4* inv_mm = nullptr;  // .w == 0: static, not in the graph
    const uint32_t* off = nullptr;   // contact CSR
    const uint32_t* peer = nullptr;
    const uint32_t* j_off = nullptr;  // joint CSR (may be null)
    const uint32_t* j_idx = nullptr;
    const uint4* j_abk = nullptr;
    // batched worlds: priorities hash a body's index within its scene, so
    // each scene is coloured as it would be alone (null: one scene)
    const uint32_t* scene_of = nullptr;
    const uint32_t* scene_off = nullptr;
};

struct ColourBuffers {
    uint32_t* colour = nullptr;    // n
    uint32_t* round_of = nullptr;  // n
    uint32_t* pending = nullptr;   // kColourRoundSlots: 1 when a body stayed uncoloured in round r; [last] overflow
    uint32_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* counts = nullptr;    // kMaxColours + 1 per sort key (the last: bodies not in the graph)
    uint32_t* sort_scratch = nullptr;
};

// Words of device memory ColourBuffers needs for n_cap bodies.
size_t colour_buffer_words(uint32_t n_cap);
void colour_buffers_bind(ColourBuffers& b, uint32_t* base, uint32_t n_cap);

// Mark every body uncoloured (static bodies: not in the graph); clear flags.
void launch_colour_init(const ColourGraph& g, const ColourBuffers& b, hipStream_t s);
// Rounds first .. first + count - 1 (each exits at once when the previous
// round left nothing uncoloured).
void launch_colour_rounds(const ColourGraph& g, const ColourBuffers& b, uint32_t first, uint32_t count,
                          hipStream_t s);
// Rounds first, first + 1, ... in one workgroup while a body stays
// uncoloured (first >= 1; exits at once when round first - 1 finished the
// colouring).  Lets a fixed batch of round launches be captured.
void launch_colour_finish(const ColourGraph& g, const ColourBuffers& b, uint32_t first, hipStream_t s);
// Sort the bodies by colour and count each colour; returns the buffer index
// (0 or 1) of keys / vals holding the sorted (colour, body) pairs.
int launch_colour_lists(const ColourGraph& g, const ColourBuffers& b, hipStream_t s);
// From the counts of the last launch_colour_lists: starts[c] .. starts[c + 1]
// is colour c's range of the sorted list, starts[kMaxColours + 1] the number
// of colours (kMaxColours + 2 words).
constexpr uint32_t kColourStartsWords = kMaxColours + 2;
void launch_colour_starts(const ColourBuffers& b, uint32_t* starts, hipStream_t s);

}  // namespace dio
