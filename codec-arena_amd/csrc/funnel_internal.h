// funnel_internal.h -- declarations shared by the host half (funnel_host.cpp)
// and the device half (screen_device.hip) of libcsfunnel.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/csfunnel.h"

namespace csf {

// Thread-local error message behind cs_last_error().
int fail(int code, const std::string &msg);
void clear_error();

// One candidate that passed a scoring launch's bound, as the scoring kernel
// writes it.  Field order keeps the struct 8-byte aligned without holes.
struct Record {
    float score;
    uint32_t stream;
    uint64_t index;  // candidate index within the stream
    uint64_t start;  // draw counter before the candidate's first draw
    uint32_t ia;     // elite index of parent a (lineage)
    uint32_t ib;     // elite index of parent b
    float g[CS_NTOT];
    uint32_t _pad;
    double score64;  // the FP64 scorer's score (the others: score)
};
static_assert(sizeof(Record) == 160, "Record layout");

// Bounded min-heap of the reference (src/screen.zig:151-193), one per
// stream, ordered by a double key: the f32 score (exactly: the conversion
// keeps order and equality) or the FP64 scorer's score.
struct HeapItem {
    double key;
    cs_survivor s;
};
struct StreamHeap {
    uint64_t k = 0;
    std::vector<HeapItem> slots;  // slots.size() == n
    void reset(uint64_t keep) {
        k = keep;
        slots.clear();
        slots.reserve(keep < (1u << 20) ? keep : (1u << 20));
    }
    bool full() const { return slots.size() >= k; }
    double root() const { return slots[0].key; }
    void offer(const cs_survivor &s, double key);
};

// Zig {d:.4} of an f32; appends to `out`.
void format_d4(float v, std::string &out);

}  // namespace csf
