// DbFile.h -- the reference's on-disk protein database and FASTA input.
//
// Three files per database, as the reference's DBmaker writes them
// (src/FastaFile.cpp:72-359, src/Bucket.cpp) and its DataPool reads them
// (src/DataPool.cpp:145-303):
//   <base>.seq    u64 offset just past the last batch, u64 0; then batches,
//                 each '$' u64 numSeqs u64 seqLen '#' and numSeqs * seqLen
//                 residue bytes, one sequence after another.  Codes 0..22 for
//                 letters, 23 for padding ('@'), 0xFF for letters without a
//                 code.  Every sequence is padded to a multiple of 4.
//   <base>.title  the '>' line of every sequence, in .seq order; none for the
//                 all-pad fillers that complete the final batch
//   <base>.map    u64 number of FASTA records, then for each run of 128
//                 consecutive .seq sequences the .title offset of the first
//                 one's line
// A sequence's index is its position in .seq, fillers included: the index the
// reference's ResultList holds and its title lookup takes.
#pragma once
#include <cstdint>
#include <fstream>
#include <mutex>
#include <string>
#include <vector>

namespace swhy {

constexpr uint8_t kUnknownCode = 0xFF;

// FASTA letter -> residue code, the reference's table (src/DataPool.cpp:33-47):
// A..Z without J, O, U in the order ABCDEFGHIKLMNPQRSTVWXYZ -> 0..22, either
// case; '@' -> 23; anything else -> kUnknownCode.
uint8_t residue_code(unsigned char c);

// Streams the records of a FASTA file: a '>' line, then sequence lines until
// the next '>' line.  Line ends ('\n', "\r\n") and blanks are dropped.
class FastaReader {
  public:
    explicit FastaReader(const std::string& path);   // throws std::runtime_error
    // title: the '>' line without its line end.  false at end of file.
    bool next(std::string& title, std::string& residues);

  private:
    std::ifstream in_;
    std::string pending_;   // a '>' line read ahead
    bool have_pending_ = false;
    std::string path_;
};

struct DbMakeStats {
    uint64_t records = 0;    // FASTA records
    uint64_t slots = 0;      // sequences in .seq, fillers included
    uint64_t batches = 0;
    uint64_t residues = 0;   // residue bytes in .seq, padding included
};

// The DBmaker: FASTA -> <base>.seq/.map/.title.  Sequences are padded to a
// multiple of 4 and bucketed by that length; whenever the residues read since
// the last flush reach flush_bytes (the reference's 256 MB), every bucket's
// whole multiples of 128 sequences go out as one batch each; at the end the
// rest is packed into 128-sequence batches of mixed lengths, every sequence
// padded at its own end to the batch's length.  Throws std::runtime_error.
DbMakeStats make_db_files(const std::string& fasta_path, const std::string& base,
                          uint64_t flush_bytes = uint64_t(1) << 28);

class DbFile {
  public:
    struct Batch {
        uint64_t first;   // index of the batch's first sequence
        uint32_t num;
        uint32_t len;
    };
    explicit DbFile(const std::string& base);   // throws std::runtime_error
    ~DbFile();
    DbFile(const DbFile&) = delete;
    DbFile& operator=(const DbFile&) = delete;

    uint64_t records() const { return records_; }
    uint64_t slots() const { return lengths_.size(); }
    uint64_t residues() const { return residues_; }   // bytes in batches, padding included
    const std::vector<Batch>& batches() const { return batches_; }
    // Per sequence: start in residue_data(), and length without its trailing
    // pad codes (fillers have length 0).
    const uint8_t* residue_data() const { return data_; }
    const uint64_t* offsets() const { return offsets_.data(); }
    const uint32_t* lengths() const { return lengths_.data(); }
    // The '>' line of sequence index (no line end); "" for fillers.
    std::string title(uint64_t index);

  private:
    struct Mapping {   // the .seq file, mapped read-only
        uint8_t* base = nullptr;
        size_t len = 0;
        ~Mapping();
    };
    Mapping seq_;
    const uint8_t* data_ = nullptr;
    uint64_t records_ = 0, residues_ = 0;
    std::vector<Batch> batches_;
    std::vector<uint64_t> offsets_;
    std::vector<uint32_t> lengths_;
    std::vector<uint64_t> title_map_;   // per 128 sequences: .title offset
    std::ifstream titles_;
    std::mutex title_mu_;
};

}  // namespace swhy
