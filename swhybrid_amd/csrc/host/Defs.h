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
// The largest chunk a worker's DataPool buffer holds.  The reference's Worker
// allocates exactly this many bytes up front (1 << 28, src/Defs.h:14,
// src/Worker.cpp:59); this pipeline's Worker allocates bufferSize bytes at the
// first fill, so its bound is only a sanity limit.
const size_t maxReaderBufferSize = size_t(16) << 30;
const int kPadCode = 23;
