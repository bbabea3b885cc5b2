#include "DbFile.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <unordered_map>

namespace swhy {

namespace {

const unsigned kBatchWidth = 128;   // sequences per batch and per map entry (src/Defs.h:15)
const uint8_t kPad = 23;

struct CodeTable {
    uint8_t t[256];
    CodeTable() {
        std::memset(t, kUnknownCode, sizeof(t));
        const char* order = "ABCDEFGHIKLMNPQRSTVWXYZ";
        for (int i = 0; order[i]; ++i) {
            t[(unsigned char)order[i]] = (uint8_t)i;
            t[(unsigned char)(order[i] - 'A' + 'a')] = (uint8_t)i;
        }
        t[(unsigned char)'@'] = kPad;
    }
};
const CodeTable kCodes;

void put_u64(std::ostream& o, uint64_t v) { o.write(reinterpret_cast<const char*>(&v), 8); }

uint64_t get_u64(const uint8_t* p) {
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}

// Sequences of one padded length waiting to be written.
struct Bucket {
    uint32_t len = 0;
    std::string seq;              // n * len residue bytes
    std::string titles;           // n title lines, each ending in '\n'
    std::vector<uint64_t> ends;   // end of each title line within titles

    uint64_t n() const { return ends.size(); }
    void add(const std::string& residues, const std::string& title) {
        seq += residues;
        titles += title;
        titles += '\n';
        ends.push_back(titles.size());
    }
    // Moves the first k sequences out, each padded at its end to out_len.
    void take(uint64_t k, uint32_t out_len, std::string& bytes, std::string& text) {
        if (k == 0) return;
        if (out_len == len) {
            bytes.append(seq, 0, k * len);
        } else {
            for (uint64_t i = 0; i < k; ++i) {
                bytes.append(seq, i * len, len);
                bytes.append(out_len - len, (char)kPad);
            }
        }
        const uint64_t cut = ends[k - 1];
        text.append(titles, 0, cut);
        seq.erase(0, k * len);
        titles.erase(0, cut);
        ends.erase(ends.begin(), ends.begin() + k);
        for (auto& e : ends) e -= cut;
    }
    // .title offsets, relative to the first title, of every 128th of the
    // first k sequences.
    std::vector<uint64_t> group_starts(uint64_t k) const {
        std::vector<uint64_t> g;
        for (uint64_t i = 0; i < k; i += kBatchWidth) g.push_back(i == 0 ? 0 : ends[i - 1]);
        return g;
    }
};

class Writer {
  public:
    explicit Writer(const std::string& base)
        : seq_(base + ".seq", std::ios::binary | std::ios::trunc),
          map_(base + ".map", std::ios::binary | std::ios::trunc),
          title_(base + ".title", std::ios::binary | std::ios::trunc) {
        if (!seq_ || !map_ || !title_)
            throw std::runtime_error("cannot create " + base + ".seq/.map/.title");
        put_u64(seq_, 0);   // end of the batches, written at finish()
        put_u64(seq_, 0);
        put_u64(map_, 0);   // number of records, written at finish()
    }

    void batch(uint64_t num, uint32_t len, const std::string& bytes, const std::string& text,
               const std::vector<uint64_t>& group_starts) {
        seq_.put('$');
        put_u64(seq_, num);
        put_u64(seq_, len);
        seq_.put('#');
        seq_.write(bytes.data(), (std::streamsize)bytes.size());
        for (uint64_t g : group_starts) put_u64(map_, title_pos_ + g);
        title_.write(text.data(), (std::streamsize)text.size());
        title_pos_ += text.size();
        st.slots += num;
        st.residues += bytes.size();
        st.batches += 1;
    }

    void finish(uint64_t records) {
        const uint64_t end = (uint64_t)seq_.tellp();
        seq_.seekp(0);
        put_u64(seq_, end);
        map_.seekp(0);
        put_u64(map_, records);
        st.records = records;
        seq_.flush();
        map_.flush();
        title_.flush();
        if (!seq_ || !map_ || !title_) throw std::runtime_error("database write failed");
    }

    DbMakeStats st;

  private:
    std::ofstream seq_, map_, title_;
    uint64_t title_pos_ = 0;
};

// Every bucket's whole multiples of 128, one batch per bucket, in bucket order
// (src/FastaFile.cpp:266-281).
void flush_full(std::vector<Bucket>& buckets, Writer& w) {
    for (Bucket& b : buckets) {
        const uint64_t k = b.n() - b.n() % kBatchWidth;
        if (k == 0) continue;
        const std::vector<uint64_t> g = b.group_starts(k);
        std::string bytes, text;
        b.take(k, b.len, bytes, text);
        w.batch(k, b.len, bytes, text, g);
    }
}

// What is left (fewer than 128 per length) in batches of exactly 128, the
// composition of src/FastaFile.cpp:293-358 with every sequence padded at its
// own end: lengths ascending; when the running count reaches 128 at bucket j,
// the batch is bucket j's share first, then every shorter sequence still
// waiting, ascending; the last batch takes the rest longest first and is
// completed with all-pad fillers.
void flush_rest(std::vector<Bucket>& buckets, Writer& w) {
    std::sort(buckets.begin(), buckets.end(),
              [](const Bucket& a, const Bucket& b) { return a.len < b.len; });
    uint64_t num = 0;
    for (size_t j = 0; j < buckets.size(); ++j) {
        if (buckets[j].n() == 0) continue;
        num += buckets[j].n();
        if (num < kBatchWidth) continue;
        const uint64_t rem = num % kBatchWidth;   // < 128: every bucket holds < 128 here
        const uint32_t L = buckets[j].len;
        std::string bytes, text;
        buckets[j].take(buckets[j].n() - rem, L, bytes, text);
        for (size_t i = 0; i < j; ++i) buckets[i].take(buckets[i].n(), L, bytes, text);
        w.batch(kBatchWidth, L, bytes, text, {0});
        num = rem;
    }
    if (num == 0) return;
    uint32_t L = 0;
    for (const Bucket& b : buckets)
        if (b.n() > 0) L = std::max(L, b.len);
    std::string bytes, text;
    for (size_t j = buckets.size(); j-- > 0;) buckets[j].take(buckets[j].n(), L, bytes, text);
    bytes.append((size_t)(kBatchWidth - num) * L, (char)kPad);
    w.batch(kBatchWidth, L, bytes, text, {0});
}

}  // namespace

uint8_t residue_code(unsigned char c) { return kCodes.t[c]; }

FastaReader::FastaReader(const std::string& path) : in_(path, std::ios::binary), path_(path) {
    if (!in_) throw std::runtime_error("cannot open " + path);
}

static void chomp(std::string& s) {
    while (!s.empty() && (s.back() == '\r' || s.back() == '\n')) s.pop_back();
}

bool FastaReader::next(std::string& title, std::string& residues) {
    title.clear();
    residues.clear();
    std::string line;
    if (!have_pending_) {
        while (std::getline(in_, line)) {
            chomp(line);
            if (line.empty()) continue;
            if (line[0] != '>') throw std::runtime_error(path_ + ": FASTA record without a '>' line");
            pending_ = line;
            have_pending_ = true;
            break;
        }
        if (!have_pending_) return false;
    }
    title = pending_;
    have_pending_ = false;
    while (std::getline(in_, line)) {
        chomp(line);
        if (!line.empty() && line[0] == '>') {
            pending_ = line;
            have_pending_ = true;
            break;
        }
        if (line.find_first_of(" \t") == std::string::npos) {
            residues += line;
        } else {
            for (char c : line)
                if (c != ' ' && c != '\t') residues.push_back(c);
        }
    }
    return true;
}

DbMakeStats make_db_files(const std::string& fasta_path, const std::string& base,
                          uint64_t flush_bytes) {
    FastaReader reader(fasta_path);
    Writer w(base);
    std::vector<Bucket> buckets;
    std::unordered_map<uint32_t, size_t> by_len;
    std::string title, letters, codes;
    uint64_t records = 0, since_flush = 0;
    while (reader.next(title, letters)) {
        // pad to a multiple of 4 with '@'; an empty record becomes 4 pads
        size_t padded = (letters.size() + 3) & ~size_t(3);
        if (padded == 0) padded = 4;
        codes.resize(padded);
        for (size_t i = 0; i < letters.size(); ++i)
            codes[i] = (char)kCodes.t[(unsigned char)letters[i]];
        for (size_t i = letters.size(); i < padded; ++i) codes[i] = (char)kPad;
        if (padded > UINT32_MAX) throw std::runtime_error("sequence too long");
        auto it = by_len.find((uint32_t)padded);
        if (it == by_len.end()) {
            it = by_len.emplace((uint32_t)padded, buckets.size()).first;
            buckets.emplace_back();
            buckets.back().len = (uint32_t)padded;
        }
        buckets[it->second].add(codes, title);
        ++records;
        since_flush += padded;
        if (since_flush + title.size() + 1 >= flush_bytes) {   // src/FastaFile.cpp:213-218
            flush_full(buckets, w);
            since_flush = 0;
        }
    }
    flush_full(buckets, w);
    flush_rest(buckets, w);
    w.finish(records);
    return w.st;
}

DbFile::Mapping::~Mapping() {
    if (base) munmap(base, len);
}

DbFile::DbFile(const std::string& base) {
    const std::string seq_path = base + ".seq";
    const int fd = ::open(seq_path.c_str(), O_RDONLY);
    if (fd < 0) throw std::runtime_error("cannot open " + seq_path);
    struct stat sb;
    if (fstat(fd, &sb) != 0 || sb.st_size < 16) {
        ::close(fd);
        throw std::runtime_error(seq_path + ": not a sequence file");
    }
    void* m = mmap(nullptr, (size_t)sb.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
    ::close(fd);
    if (m == MAP_FAILED) throw std::runtime_error("cannot map " + seq_path);
    seq_.base = static_cast<uint8_t*>(m);
    seq_.len = (size_t)sb.st_size;
    const uint8_t* f = seq_.base;
    data_ = f;
    const uint64_t end = get_u64(f);
    if (end > seq_.len || end < 16) throw std::runtime_error(seq_path + ": bad batch end offset");
    uint64_t p = 16;
    while (p < end) {
        if (p + 18 > end || f[p] != '$' || f[p + 17] != '#')
            throw std::runtime_error(seq_path + ": damaged batch header");
        const uint64_t num = get_u64(f + p + 1);
        const uint64_t len = get_u64(f + p + 9);
        p += 18;
        if (num > UINT32_MAX || len > UINT32_MAX || (len > 0 && num > (end - p) / len))
            throw std::runtime_error(seq_path + ": batch larger than the file");
        batches_.push_back(Batch{(uint64_t)lengths_.size(), (uint32_t)num, (uint32_t)len});
        for (uint64_t k = 0; k < num; ++k) {
            const uint64_t off = p + k * len;
            uint32_t n = (uint32_t)len;
            while (n > 0 && f[off + n - 1] == kPad) --n;
            offsets_.push_back(off);
            lengths_.push_back(n);
        }
        p += num * len;
        residues_ += num * len;
    }
    std::ifstream mf(base + ".map", std::ios::binary);
    if (!mf) throw std::runtime_error("cannot open " + base + ".map");
    uint64_t v;
    if (!mf.read(reinterpret_cast<char*>(&v), 8)) throw std::runtime_error(base + ".map: empty");
    records_ = v;
    while (mf.read(reinterpret_cast<char*>(&v), 8)) title_map_.push_back(v);
    titles_.open(base + ".title", std::ios::binary);
    if (!titles_) throw std::runtime_error("cannot open " + base + ".title");
}

DbFile::~DbFile() = default;

// src/DataPool.cpp:289-303: the map entry of the index's run of 128, then
// that many lines on.  Fillers (indices past the records) have no title.
std::string DbFile::title(uint64_t index) {
    if (index >= records_ || index >= slots()) return std::string();
    const uint64_t g = index / kBatchWidth;
    if (g >= title_map_.size()) return std::string();
    std::lock_guard<std::mutex> lk(title_mu_);
    titles_.clear();
    titles_.seekg((std::streamoff)title_map_[g]);
    std::string line;
    for (uint64_t i = 0; i <= index % kBatchWidth; ++i)
        if (!std::getline(titles_, line)) return std::string();
    chomp(line);
    return line;
}

}  // namespace swhy
