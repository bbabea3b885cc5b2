// Defs.h -- shared host-side definitions of the worker pipeline.
#pragma once
#include <cstddef>
#include <cstdint>

// One batch of a database chunk: numSeqs sequences of seqLen residues each,
// stored back to back (the record of src/Defs.h:24-29 / src/DataPool.cpp:171-278).
struct batchInfo {
    batchInfo() : numSeqs(0), seqLen(0) {}
    batchInfo(unsigned n, unsigned l) : numSeqs(n), seqLen(l) {}
    unsigned numSeqs;
    unsigned seqLen;
};

const int alphabetSize = 32;
const int kPadCode = 23;
