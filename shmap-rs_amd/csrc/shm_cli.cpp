// shmap -- command-line front end over the gfx950 mapping library.
//
// Keeps the reference CLI (src/main.rs:16-70, src/params.rs:14-263): the same
// flags, defaults and validation messages, PAF on stdout in input order,
// <-z>.unmapped.paf for unmapped reads and --per-read-stats rows
// (src/shmap/mod.rs:280-327, 374-650).  Reads are mapped on the GPU in batches;
// records are written strictly in input order.
#include <errno.h>
#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <unistd.h>
#include <zlib.h>

#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/shmap_hip.h"

namespace {

struct Params {
    std::string p_file, t_file, params_file, per_read_stats, profile_log;
    int32_t k = 15;
    double h_frac = 0.05, theta = 0.9, min_diff = 0.02, max_overlap = 0.5;
    double rarity_weight = 0, rarity_tiebreak = 0, rarity_alpha = 0;
    bool have_max_seeds = false, have_max_matches = false;
    int32_t max_seeds = 0, max_matches = 0;
    int32_t metric = SHM_CONTAINMENT;
    int32_t verbose = 0;
    bool normalize = false, nbp = false, one_sweep = false, abs_pos = false, profile = false;
    long long threads = 1;
    unsigned long long per_read_stats_sample = 1;
    int device = 0;
    long batch = 16384;
    bool dump_reads = false;
    std::string save_index;
};

const char *metric_name(int32_t m) {
    switch (m) {
    case SHM_JACCARD: return "Jaccard";
    case SHM_BUCKET_SH: return "bucket_SH";
    case SHM_BUCKET_LCS: return "bucket_LCS";
    default: return "Containment";
    }
}

// Rust's Display for f64: the shortest decimal that round-trips, never in
// exponent form.
std::string rust_f64(double v) {
    char buf[64];
    for (int p = 1; p <= 17; p++) {
        snprintf(buf, sizeof buf, "%.*e", p - 1, v);
        if (strtod(buf, nullptr) == v) break;
    }
    // buf holds d.ddde[+-]xx; expand to plain decimal
    char mant[32];
    int ex = 0;
    bool neg = buf[0] == '-';
    const char *s = buf + (neg ? 1 : 0);
    int mi = 0;
    for (; *s && *s != 'e'; s++)
        if (*s != '.') mant[mi++] = *s;
    mant[mi] = 0;
    if (*s == 'e') ex = atoi(s + 1);
    while (mi > 1 && mant[mi - 1] == '0') mant[--mi] = 0;
    std::string out = neg ? "-" : "";
    const int point = ex + 1;  // digits before the decimal point
    if (point <= 0) {
        out += "0.";
        out += std::string(-point, '0');
        out += mant;
    } else if (point >= mi) {
        out += mant;
        out += std::string(point - mi, '0');
    } else {
        out += std::string(mant, point);
        out += ".";
        out += mant + point;
    }
    return out;
}

// True when path holds an index written by --save-index (its magic bytes).
static bool is_index_file(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    char m[8];
    const bool yes = fread(m, 1, 8, f) == 8 && memcmp(m, "SHMIDX01", 8) == 0;
    fclose(f);
    return yes;
}

void usage(FILE *f) {
    fprintf(f,
            "Map-SHmap: fast and accurate sketch-based read mapper (MI355X)\n\n"
            "Usage: shmap -p <P_FILE> -s <T_FILE> [OPTIONS]\n\n"
            "  -p, --pattern <P_FILE>        Reads file (FASTA format)\n"
            "  -s, --text <T_FILE>           Reference file (FASTA format), or an index written by --save-index\n"
            "  -k, --ksize <K>               K-mer length [default: 15]\n"
            "  -r, --hashratio <H_FRAC>      FracMinHash ratio in (0,1] [default: 0.05]\n"
            "  -S, --max_seeds <MAX_SEEDS>   Maximum seeds in a sketch (accepted, unused)\n"
            "  -M, --max_matches <M>         Maximum seed matches in a sketch\n"
            "  -t, --threshold <THETA>       Homology threshold in [0,1] [default: 0.9]\n"
            "  -d, --min_diff <MIN_DIFF>     Minimum difference between best and second best [default: 0.02]\n"
            "  -o, --max_overlap <O>         Maximum overlap between best and second best [default: 0.5]\n"
            "  -m, --metric <METRIC>         Containment, Jaccard, bucket_SH, bucket_LCS [default: Containment]\n"
            "  -z, --params <PARAMS_FILE>    Output file with parameters (tsv)\n"
            "  -v, --verbose <VERBOSE>       Verbosity level 0 or 1 [default: 0]\n"
            "  -n, --normalize               Accepted, unused\n"
            "  -P, --no_bucket_pruning       Disables bucket pruning\n"
            "  -B, --one_sweep               One sweep on all matches\n"
            "  -F, --abs_pos                 Absolute positions instead of k-mer positions\n"
            "  -@, --threads <THREADS>       Host threads reading FASTA and writing records (mapping runs on the GPU) [default: 1]\n"
            "  -x, --profile                 Write a timing report (JSON) to --profile-log\n"
            "      --profile-log <FILE>\n"
            "      --per-read-stats <FILE>   One TSV row per read\n"
            "      --per-read-stats-sample <N>  Keep every Nth read [default: 1]\n"
            "      --device <ID>             GPU device [default: 0]\n"
            "      --batch <N>               Reads per GPU batch [default: 16384]\n"
            "      --save-index <FILE>       Write the index built from -s to FILE (reused as -s FILE)\n"
            "      --dump-reads              Print the reads of -p as parsed (FASTA) and exit\n");
}

struct Record {
    std::string name;
    std::vector<uint8_t> seq;
};

// A batch of reads laid out contiguously: read i is bases[off[i], off[i+1])
// with id names[name_off[i], name_off[i+1]).
struct Batch {
    std::vector<uint8_t> bases;
    std::vector<int64_t> off{0};
    std::string names;
    std::vector<size_t> name_off{0};
    size_t size() const { return off.size() - 1; }
    void clear() {
        bases.clear();
        off.assign(1, 0);
        names.clear();
        name_off.assign(1, 0);
    }
    std::string name(size_t i) const { return names.substr(name_off[i], name_off[i + 1] - name_off[i]); }
};

// FASTA/FASTQ records as needletail delivers them (src/io/mod.rs:383-427): the
// id cut at the first space or tab, line breaks and '\r' removed.  The input
// is read in large blocks and split into lines with memchr; a record's bases
// are copied straight into the caller's storage.  Plain and gzip input (by
// its magic bytes, any number of members) are read; bzip2, xz and zstd input,
// which needletail also reads, is refused with a message.
class FastxReader {
  public:
    explicit FastxReader(const std::string &path) : buf_(1u << 24) {
        FILE *f = fopen(path.c_str(), "rb");
        if (!f) {
            err_ = "failed to open \"" + path + "\"";
            return;
        }
        unsigned char m[6] = {0, 0, 0, 0, 0, 0};
        const size_t got = fread(m, 1, sizeof m, f);
        const char *fmt = nullptr;
        if (got >= 3 && m[0] == 'B' && m[1] == 'Z' && m[2] == 'h') fmt = "bzip2";
        else if (got >= 6 && m[0] == 0xfd && m[1] == '7' && m[2] == 'z' && m[3] == 'X' && m[4] == 'Z' && m[5] == 0) fmt = "xz";
        else if (got >= 4 && m[0] == 0x28 && m[1] == 0xb5 && m[2] == 0x2f && m[3] == 0xfd) fmt = "zstd";
        if (fmt) {
            fclose(f);
            err_ = std::string("\"") + path + "\" is " + fmt +
                   "-compressed; this build reads plain or gzip FASTA/FASTQ (decompress it first)";
            return;
        }
        if (got >= 2 && m[0] == 0x1f && m[1] == 0x8b) {
            fclose(f);
            gz_ = gzopen(path.c_str(), "rb");
            if (!gz_) {
                err_ = "failed to open \"" + path + "\"";
                return;
            }
            gzbuffer(gz_, 1u << 20);
        } else {
            rewind(f);
            f_ = f;
        }
    }
    ~FastxReader() {
        if (f_) fclose(f_);
        if (gz_) gzclose(gz_);
    }
    bool ok() const { return f_ != nullptr || gz_ != nullptr; }
    // why the input could not be opened or read ("" if it was read whole)
    const std::string &error() const { return err_; }
    bool next(Record &r) {
        r.name.clear();
        r.seq.clear();
        std::vector<uint8_t> *dst = &r.seq;
        return record(
            [&](const char *s, size_t n) { r.name.assign(s, n); },
            [&](const char *s, size_t n) { dst->insert(dst->end(), (const uint8_t *)s, (const uint8_t *)s + n); });
    }
    // Appends up to max_reads records; false when none was left.
    bool next_batch(Batch &b, long max_reads) {
        b.clear();
        while ((long)b.size() < max_reads) {
            const bool got = record([&](const char *s, size_t n) { b.names.append(s, n); },
                                    [&](const char *s, size_t n) { b.bases.insert(b.bases.end(), s, s + n); });
            if (!got) break;
            b.off.push_back((int64_t)b.bases.size());
            b.name_off.push_back(b.names.size());
        }
        return b.size() > 0;
    }

  private:
    // The next line without its '\n' (and a trailing '\r'), valid until the
    // next call; false at the end of the input.
    bool line(const char *&s, size_t &n) {
        for (;;) {
            const char *p = (const char *)memchr(buf_.data() + pos_, '\n', end_ - pos_);
            if (p || eof_) {
                if (!p && pos_ == end_) return false;
                const size_t e = p ? (size_t)(p - buf_.data()) : end_;
                s = buf_.data() + pos_;
                n = e - pos_;
                pos_ = p ? e + 1 : end_;
                if (n && s[n - 1] == '\r') n--;
                return true;
            }
            // no complete line left: keep the tail, grow if one line fills the buffer, read more
            memmove(buf_.data(), buf_.data() + pos_, end_ - pos_);
            end_ -= pos_;
            pos_ = 0;
            if (end_ == buf_.size()) buf_.resize(buf_.size() * 2);
            const size_t got = fill(buf_.data() + end_, buf_.size() - end_);
            end_ += got;
            if (got == 0) eof_ = true;
        }
    }
    template <typename NameF, typename SeqF>
    bool record(NameF name, SeqF seq) {
        const char *s;
        size_t n;
        if (have_pending_) {
            have_pending_ = false;
            s = pending_.data();
            n = pending_.size();
        } else {
            do {
                if (!line(s, n)) return false;
            } while (n == 0 || (s[0] != '>' && s[0] != '@'));
        }
        const bool fastq = s[0] == '@';
        size_t e = 1;
        while (e < n && s[e] != ' ' && s[e] != '\t') e++;
        name(s + 1, e - 1);
        if (fastq) {
            if (line(s, n)) seq(s, n);
            line(s, n);  // '+'
            line(s, n);  // qualities
            return true;
        }
        while (line(s, n)) {
            if (n && s[0] == '>') {
                pending_.assign(s, n);
                have_pending_ = true;
                break;
            }
            seq(s, n);
        }
        return true;
    }
    size_t fill(char *dst, size_t n) {
        if (f_) return fread(dst, 1, n, f_);
        if (!gz_ || !err_.empty()) return 0;
        const int got = gzread(gz_, dst, (unsigned)(n < (1u << 30) ? n : (1u << 30)));
        if (got < 0) {
            int code = 0;
            const char *msg = gzerror(gz_, &code);
            err_ = std::string("failed to decompress the input: ") + (msg ? msg : "gzip error");
            return 0;
        }
        return (size_t)got;
    }
    FILE *f_ = nullptr;
    gzFile gz_ = nullptr;
    std::string err_;
    std::vector<char> buf_;
    size_t pos_ = 0, end_ = 0;
    bool eof_ = false;
    std::string pending_;
    bool have_pending_ = false;
};

double now_s() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

// Plain FASTA read in parallel byte ranges, as the reference's two-pass reader
// does (src/io/mod.rs:198-355): the file is memory-mapped and taken a window
// at a time, each window cut at record starts (a '>' at the file's start or
// right after a '\n'), and the window's bytes into one range per thread; a
// thread parses the records that start in its range, reading past the range's
// end to finish its last one.  Records come out in file order, one batch per
// range, with FastxReader's semantics (id cut at the first space or tab,
// '\r' before a line break dropped, text before the first record skipped).
class ParallelFasta {
  public:
    ParallelFasta(const std::string &path, int threads, size_t window)
        : threads_(threads < 1 ? 1 : threads), window_(window < (1u << 20) ? (1u << 20) : window) {
        fd_ = open(path.c_str(), O_RDONLY);
        if (fd_ < 0) {
            err_ = "failed to open \"" + path + "\"";
            return;
        }
        struct stat st;
        if (fstat(fd_, &st) != 0) {
            err_ = "failed to read \"" + path + "\"";
            return;
        }
        size_ = (size_t)st.st_size;
        if (size_ > 0) {
            void *m = mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd_, 0);
            if (m == MAP_FAILED) {
                err_ = "failed to map \"" + path + "\"";
                return;
            }
            data_ = (const char *)m;
            madvise(m, size_, MADV_SEQUENTIAL);
        }
        pos_ = start_at(0);
        ok_ = true;
    }
    ~ParallelFasta() {
        if (data_) munmap((void *)data_, size_);
        if (fd_ >= 0) close(fd_);
    }
    ParallelFasta(const ParallelFasta &) = delete;
    ParallelFasta &operator=(const ParallelFasta &) = delete;
    bool ok() const { return ok_; }
    const std::string &error() const { return err_; }

    // Plain (uncompressed) input whose first record is FASTA; FASTQ and
    // compressed input take FastxReader.
    static bool usable(const std::string &path) {
        FILE *f = fopen(path.c_str(), "rb");
        if (!f) return false;
        char buf[1 << 16];
        const size_t n = fread(buf, 1, sizeof buf, f);
        fclose(f);
        if (n >= 2 && (unsigned char)buf[0] == 0x1f && (unsigned char)buf[1] == 0x8b) return false;
        if (n >= 3 && buf[0] == 'B' && buf[1] == 'Z' && buf[2] == 'h') return false;
        if (n >= 4 && (unsigned char)buf[0] == 0x28 && (unsigned char)buf[1] == 0xb5) return false;
        if (n >= 6 && (unsigned char)buf[0] == 0xfd && buf[1] == '7') return false;
        // the first line starting a record decides (FastxReader::record)
        for (size_t i = 0; i < n; i++)
            if (i == 0 || buf[i - 1] == '\n') {
                if (buf[i] == '>') return true;
                if (buf[i] == '@') return false;
            }
        return false;
    }

    // The next window's records, one batch per thread range, in order; false
    // when the file is done.
    bool next_window(std::vector<Batch> &out) {
        if (!ok_ || pos_ >= size_) return false;
        const size_t we = start_at(pos_ + window_ < size_ ? pos_ + window_ : size_);
        out.resize(threads_);
        std::vector<std::thread> th;
        for (int t = 0; t < threads_; t++) {
            const size_t a = pos_ + (we - pos_) * t / threads_, b = pos_ + (we - pos_) * (t + 1) / threads_;
            th.emplace_back([this, a, b, t, &out] { parse(a, b, out[t]); });
        }
        for (std::thread &x : th) x.join();
        pos_ = we;
        return true;
    }

  private:
    // The first record start at or after p (or the file's size).
    size_t start_at(size_t p) const {
        if (p == 0 && size_ > 0 && data_[0] == '>') return 0;
        size_t q = p > 0 ? p - 1 : 0;
        while (q < size_) {
            const char *nl = (const char *)memchr(data_ + q, '\n', size_ - q);
            if (!nl) return size_;
            q = (size_t)(nl - data_) + 1;
            if (q < size_ && data_[q] == '>') return q;
        }
        return size_;
    }
    // The line at p: [p, e) without its '\n' and a '\r' before it; the next
    // line's start in *next.
    void line_at(size_t p, size_t &e, size_t &next) const {
        const char *nl = (const char *)memchr(data_ + p, '\n', size_ - p);
        const size_t end = nl ? (size_t)(nl - data_) : size_;
        next = nl ? end + 1 : size_;
        e = (end > p && data_[end - 1] == '\r') ? end - 1 : end;
    }
    void parse(size_t a, size_t b, Batch &out) const {
        out.clear();
        size_t p = start_at(a);
        while (p < b) {
            size_t e, next;
            line_at(p, e, next);
            size_t ne = p + 1;
            while (ne < e && data_[ne] != ' ' && data_[ne] != '\t') ne++;
            out.names.append(data_ + p + 1, ne - p - 1);
            p = next;
            while (p < size_ && data_[p] != '>') {
                line_at(p, e, next);
                out.bases.insert(out.bases.end(), data_ + p, data_ + e);
                p = next;
            }
            out.off.push_back((int64_t)out.bases.size());
            out.name_off.push_back(out.names.size());
        }
    }
    int threads_;
    size_t window_;
    int fd_ = -1;
    const char *data_ = nullptr;
    size_t size_ = 0, pos_ = 0;
    bool ok_ = false;
    std::string err_;
};

// A reads file as batches: ParallelFasta for plain FASTA when more than one
// host thread is asked for, FastxReader otherwise.
class ReadSource {
  public:
    ReadSource(const std::string &path, int threads, long max_reads) : max_(max_reads) {
        if (threads > 1 && ParallelFasta::usable(path)) pf_.reset(new ParallelFasta(path, threads, 512u << 20));
        else fx_.reset(new FastxReader(path));
    }
    bool ok() const { return pf_ ? pf_->ok() : fx_->ok(); }
    const std::string &error() const { return pf_ ? pf_->error() : fx_->error(); }
    // The next non-empty batch; false at the end.
    bool next(Batch &b) {
        if (fx_) return fx_->next_batch(b, max_);
        for (;;) {
            while (at_ < win_.size()) {
                Batch &w = win_[at_++];
                if (w.size()) {
                    std::swap(b, w);
                    return true;
                }
            }
            if (!pf_->next_window(win_)) return false;
            at_ = 0;
        }
    }

  private:
    long max_;
    std::unique_ptr<FastxReader> fx_;
    std::unique_ptr<ParallelFasta> pf_;
    std::vector<Batch> win_;
    size_t at_ = 0;
};

// Batches read ahead on a thread of their own (two in flight), so parsing
// overlaps mapping and output.
class BatchReader {
  public:
    explicit BatchReader(std::function<bool(Batch &)> src) : src_(std::move(src)), th_([this] { run(); }) {}
    ~BatchReader() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
    }
    // The next batch (nullptr at the end); give it back with recycle().
    Batch *next() {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !full_.empty() || done_; });
        if (full_.empty()) return nullptr;
        Batch *b = full_.front();
        full_.pop_front();
        cv_.notify_all();
        return b;
    }
    // Seconds the reader has spent in the source so far.
    double busy_s() {
        std::lock_guard<std::mutex> lk(mu_);
        return busy_s_;
    }
    void recycle(Batch *b) {
        std::lock_guard<std::mutex> lk(mu_);
        free_.push_back(b);
        cv_.notify_all();
    }

  private:
    void run() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (int i = 0; i < 2; i++) free_.push_back(&pool_[i]);
        }
        for (;;) {
            Batch *b;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return !free_.empty() || stop_; });
                if (stop_) return;
                b = free_.front();
                free_.pop_front();
            }
            const double t0 = now_s();
            const bool got = src_(*b);
            const double dt = now_s() - t0;
            std::lock_guard<std::mutex> lk(mu_);
            busy_s_ += dt;
            if (!got) {
                done_ = true;
                cv_.notify_all();
                return;
            }
            full_.push_back(b);
            cv_.notify_all();
        }
    }
    std::function<bool(Batch &)> src_;
    double busy_s_ = 0;  // time spent reading and parsing
    Batch pool_[2];
    std::deque<Batch *> free_, full_;
    std::mutex mu_;
    std::condition_variable cv_;
    bool stop_ = false, done_ = false;
    std::thread th_;
};


bool parse_i32(const char *s, int32_t &v) {
    char *e;
    errno = 0;
    long long x = strtoll(s, &e, 10);
    if (errno || *e || e == s || x < INT32_MIN || x > INT32_MAX) return false;
    v = (int32_t)x;
    return true;
}
bool parse_f64(const char *s, double &v) {
    char *e;
    v = strtod(s, &e);
    return !(*e || e == s);
}

int bad_value(const char *flag, const char *v) {
    fprintf(stderr, "error: invalid value '%s' for '%s'\n\nFor more information, try '--help'.\n", v, flag);
    return 2;
}

// src/params.rs:191-263 -- Params::validate
std::string validate(const Params &p) {
    char b[512];
    if (p.k <= 0) return snprintf(b, sizeof b, "K-mer length (-k) should be positive. You provided %d.", p.k), b;
    if (p.h_frac <= 0.0 || p.h_frac > 1.0)
        return snprintf(b, sizeof b, "Given hash ratio (-r) should be between 0 and 1. You provided %s.",
                        rust_f64(p.h_frac).c_str()),
               b;
    if (p.have_max_seeds && p.max_seeds <= 0)
        return snprintf(b, sizeof b, "The number of seeds (-S) should be positive. You provided %d.", p.max_seeds), b;
    if (p.have_max_matches && p.max_matches <= 0)
        return snprintf(b, sizeof b, "The number of seed matches (-M) should be positive. You provided %d.",
                        p.max_matches),
               b;
    if (p.theta < 0.0 || p.theta > 1.0)
        return snprintf(b, sizeof b, "The threshold (-t) should be between 0 and 1. You provided %s.",
                        rust_f64(p.theta).c_str()),
               b;
    if (p.min_diff < 0.0)
        return snprintf(b, sizeof b, "The minimum difference (-d) should be >= 0. You provided %s.",
                        rust_f64(p.min_diff).c_str()),
               b;
    if (p.rarity_weight < 0.0 || p.rarity_weight > 4.0)
        return snprintf(b, sizeof b, "The rarity weight (--rarity-weight) should be between 0 and 4. You provided %s.",
                        rust_f64(p.rarity_weight).c_str()),
               b;
    if (p.rarity_weight > 0.0 && p.metric != SHM_CONTAINMENT)
        return snprintf(b, sizeof b,
                        "--rarity-weight is only defined for -m Containment (got %s). Jaccard's denominator "
                        "includes reference k-mers absent from the read, which have no frequency to weight by.",
                        metric_name(p.metric)),
               b;
    if (p.max_overlap < 0.0 || p.max_overlap > 1.0)
        return snprintf(b, sizeof b, "The maximum overlap (-o) should be between 0 and 1. You provided %s.",
                        rust_f64(p.max_overlap).c_str()),
               b;
    if (p.verbose < 0 || p.verbose > 2)
        return snprintf(b, sizeof b, "--verbose (-v) should be 0, 1, or 2. You provided %d.", p.verbose), b;
    if (p.threads == 0)
        return snprintf(b, sizeof b, "The number of threads (-@) should be positive. You provided %lld.", p.threads),
               b;
    if (p.per_read_stats_sample == 0)
        return "--per-read-stats-sample should be >= 1 (1 keeps every read). You provided 0.";
    // not part of the GPU path
    if (p.rarity_weight > 0.0 || p.rarity_tiebreak > 0.0 || p.rarity_alpha > 0.0)
        return "--rarity-weight/--rarity-tiebreak/--rarity-alpha are not supported by the GPU mapper.";
    if (p.verbose >= 2) return "--verbose 2 (ground-truth analysis) is not supported by the GPU mapper.";
    return "";
}

// src/params.rs:266-304 -- the -z parameters TSV (header line, values line, no
// trailing newline after the values).
void write_params(const Params &p) {
    FILE *f = fopen(p.params_file.c_str(), "w");
    if (!f) return;
    const std::string opt_s = p.have_max_seeds ? std::to_string(p.max_seeds) : "-1";
    const std::string opt_m = p.have_max_matches ? std::to_string(p.max_matches) : "-1";
    const std::vector<std::pair<std::string, std::string>> fields = {
        {"pFile", p.p_file}, {"tFile", p.t_file}, {"k", std::to_string(p.k)}, {"hFrac", rust_f64(p.h_frac)},
        {"max_seeds", opt_s}, {"max_matches", opt_m}, {"tThres", rust_f64(p.theta)},
        {"min_diff", rust_f64(p.min_diff)}, {"max_overlap", rust_f64(p.max_overlap)}, {"paramsFile", p.params_file},
        {"metric", metric_name(p.metric)}, {"normalize", std::to_string((int)p.normalize)},
        {"verbose", std::to_string(p.verbose)}, {"no-bucket-pruning", std::to_string((int)p.nbp)},
        {"one-sweep", std::to_string((int)p.one_sweep)}, {"abs-pos", std::to_string((int)p.abs_pos)},
        {"threads", std::to_string(p.threads)}};
    for (auto &kv : fields) fprintf(f, "%s\t", kv.first.c_str());
    fprintf(f, "\n");
    for (auto &kv : fields) fprintf(f, "%s\t", kv.second.c_str());
    fclose(f);
}

}  // namespace

int main(int argc, char **argv) {
    Params p;
    bool have_p = false, have_s = false;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        std::string val;
        bool has_inline = false;
        if (a.rfind("--", 0) == 0) {
            size_t eq = a.find('=');
            if (eq != std::string::npos) {
                val = a.substr(eq + 1);
                a = a.substr(0, eq);
                has_inline = true;
            }
        }
        auto need = [&](const char *flag) -> const char * {
            if (has_inline) return val.c_str();
            if (i + 1 >= argc) {
                fprintf(stderr, "error: a value is required for '%s' but none was supplied\n", flag);
                exit(2);
            }
            return argv[++i];
        };
        if (a == "-h" || a == "--help") {
            usage(stdout);
            return 0;
        } else if (a == "-V" || a == "--version") {
            printf("shmap 1.4.3 (gfx950)\n");
            return 0;
        } else if (a == "-p" || a == "--pattern") {
            p.p_file = need("--pattern <P_FILE>");
            have_p = true;
        } else if (a == "-s" || a == "--text") {
            p.t_file = need("--text <T_FILE>");
            have_s = true;
        } else if (a == "-k" || a == "--ksize") {
            const char *v = need("--ksize <K>");
            if (!parse_i32(v, p.k)) return bad_value("--ksize <K>", v);
        } else if (a == "-r" || a == "--hashratio") {
            const char *v = need("--hashratio <H_FRAC>");
            if (!parse_f64(v, p.h_frac)) return bad_value("--hashratio <H_FRAC>", v);
        } else if (a == "-S" || a == "--max_seeds") {
            const char *v = need("--max_seeds <MAX_SEEDS>");
            if (!parse_i32(v, p.max_seeds)) return bad_value("--max_seeds <MAX_SEEDS>", v);
            p.have_max_seeds = true;
        } else if (a == "-M" || a == "--max_matches") {
            const char *v = need("--max_matches <MAX_MATCHES>");
            if (!parse_i32(v, p.max_matches)) return bad_value("--max_matches <MAX_MATCHES>", v);
            p.have_max_matches = true;
        } else if (a == "-t" || a == "--threshold") {
            const char *v = need("--threshold <THETA>");
            if (!parse_f64(v, p.theta)) return bad_value("--threshold <THETA>", v);
        } else if (a == "-d" || a == "--min_diff") {
            const char *v = need("--min_diff <MIN_DIFF>");
            if (!parse_f64(v, p.min_diff)) return bad_value("--min_diff <MIN_DIFF>", v);
        } else if (a == "-o" || a == "--max_overlap") {
            const char *v = need("--max_overlap <MAX_OVERLAP>");
            if (!parse_f64(v, p.max_overlap)) return bad_value("--max_overlap <MAX_OVERLAP>", v);
        } else if (a == "--rarity-weight") {
            const char *v = need("--rarity-weight <RARITY_WEIGHT>");
            if (!parse_f64(v, p.rarity_weight)) return bad_value("--rarity-weight <RARITY_WEIGHT>", v);
        } else if (a == "--rarity-tiebreak") {
            const char *v = need("--rarity-tiebreak <RARITY_TIEBREAK>");
            if (!parse_f64(v, p.rarity_tiebreak)) return bad_value("--rarity-tiebreak <RARITY_TIEBREAK>", v);
        } else if (a == "--rarity-alpha") {
            const char *v = need("--rarity-alpha <RARITY_ALPHA>");
            if (!parse_f64(v, p.rarity_alpha)) return bad_value("--rarity-alpha <RARITY_ALPHA>", v);
        } else if (a == "-m" || a == "--metric") {
            const char *v = need("--metric <METRIC>");
            if (!strcmp(v, "Containment")) p.metric = SHM_CONTAINMENT;
            else if (!strcmp(v, "Jaccard")) p.metric = SHM_JACCARD;
            else if (!strcmp(v, "bucket_SH")) p.metric = SHM_BUCKET_SH;
            else if (!strcmp(v, "bucket_LCS")) p.metric = SHM_BUCKET_LCS;
            else {
                fprintf(stderr,
                        "error: invalid value '%s' for '--metric <METRIC>'\n  [possible values: Containment, "
                        "Jaccard, bucket_SH, bucket_LCS]\n\nFor more information, try '--help'.\n",
                        v);
                return 2;
            }
        } else if (a == "-z" || a == "--params") {
            p.params_file = need("--params <PARAMS_FILE>");
        } else if (a == "-v" || a == "--verbose") {
            const char *v = need("--verbose <VERBOSE>");
            if (!parse_i32(v, p.verbose)) return bad_value("--verbose <VERBOSE>", v);
        } else if (a == "-n" || a == "--normalize") {
            p.normalize = true;
        } else if (a == "-P" || a == "--no_bucket_pruning") {
            p.nbp = true;
        } else if (a == "-B" || a == "--one_sweep") {
            p.one_sweep = true;
        } else if (a == "-F" || a == "--abs_pos") {
            p.abs_pos = true;
        } else if (a == "-@" || a == "--threads") {
            const char *v = need("--threads <THREADS>");
            char *e;
            p.threads = strtoll(v, &e, 10);
            if (*e || e == v || p.threads < 0) return bad_value("--threads <THREADS>", v);
        } else if (a == "-x" || a == "--profile") {
            p.profile = true;
        } else if (a == "--profile-log") {
            p.profile_log = need("--profile-log <PROFILE_LOG>");
        } else if (a == "--per-read-stats") {
            p.per_read_stats = need("--per-read-stats <PER_READ_STATS>");
        } else if (a == "--per-read-stats-sample") {
            const char *v = need("--per-read-stats-sample <PER_READ_STATS_SAMPLE>");
            char *e;
            p.per_read_stats_sample = strtoull(v, &e, 10);
            if (*e || e == v) return bad_value("--per-read-stats-sample <PER_READ_STATS_SAMPLE>", v);
        } else if (a == "--device") {
            p.device = atoi(need("--device <ID>"));
        } else if (a == "--batch") {
            p.batch = atol(need("--batch <N>"));
            if (p.batch < 1) p.batch = 1;
        } else if (a == "--dump-reads") {
            p.dump_reads = true;
        } else if (a == "--save-index") {
            p.save_index = need("--save-index <FILE>");
        } else {
            fprintf(stderr, "error: unexpected argument '%s' found\n\nFor more information, try '--help'.\n",
                    argv[i]);
            return 2;
        }
    }
    if (p.dump_reads && have_p) {
        // the reads as parsed, one FASTA record per line pair (no device needed)
        ReadSource rd(p.p_file, (int)p.threads, p.batch);
        if (!rd.ok()) {
            fprintf(stderr, "Error: %s\n", rd.error().c_str());
            return 1;
        }
        Batch b;
        while (rd.next(b))
            for (size_t i = 0; i < b.size(); i++) {
                const std::string nm = b.name(i);
                fprintf(stdout, ">%s\n", nm.c_str());
                fwrite(b.bases.data() + b.off[i], 1, (size_t)(b.off[i + 1] - b.off[i]), stdout);
                fputc('\n', stdout);
            }
        if (!rd.error().empty()) {
            fprintf(stderr, "Error: %s\n", rd.error().c_str());
            return 1;
        }
        return 0;
    }
    if (!have_p || !have_s) {
        fprintf(stderr, "error: the following required arguments were not provided:\n%s%s\nUsage: shmap -p <P_FILE> "
                        "-s <T_FILE>\n\nFor more information, try '--help'.\n",
                have_p ? "" : "  --pattern <P_FILE>\n", have_s ? "" : "  --text <T_FILE>\n");
        return 2;
    }
    const std::string err = validate(p);
    if (!err.empty()) {
        fprintf(stderr, "ERROR: %s\n", err.c_str());
        return 1;
    }
    if (!p.params_file.empty()) {
        fprintf(stderr, "Writing parameters to %s...\n", p.params_file.c_str());
        write_params(p);
    }
    const double t_start = now_s();

    // ---- index (src/main.rs:45-56)
    fprintf(stderr, "Indexing %s...\n", p.t_file.c_str());
    shm_sketcher *sk = nullptr;
    shm_index *idx = nullptr;
    const double t_idx = now_s();
    double index_s = 0;
    long long nucl = 0;
    size_t nsegs = 0;
    if (is_index_file(p.t_file)) {
        // an index written by --save-index: the build parameters must match
        if (shm_index_load(p.t_file.c_str(), p.device, &idx)) {
            fprintf(stderr, "Error: failed to load index from %s\n\nCaused by:\n    %s\n", p.t_file.c_str(),
                    shm_last_error());
            return 1;
        }
        int32_t ik = 0, im = 0;
        double ir = 0;
        shm_index_params(idx, &ik, &ir, &im);
        const int32_t want_m = p.have_max_matches ? p.max_matches : 0;
        if (ik != p.k || ir != p.h_frac || im != want_m) {
            fprintf(stderr, "ERROR: the index in %s was built with -k %d -r %g -M %d; this run asks for -k %d -r %g -M %d\n",
                    p.t_file.c_str(), ik, ir, im, p.k, p.h_frac, want_m);
            shm_index_free(idx);
            return 1;
        }
        int32_t n = 0;
        shm_index_n_segments(idx, &n);
        nsegs = (size_t)n;
        for (int32_t s = 0; s < n; s++) {
            const char *nm;
            int32_t sz;
            int64_t nk;
            shm_index_segment(idx, s, &nm, &sz, &nk);
            nucl += sz;
        }
    } else {
        std::vector<Batch> ref;
        {
            ReadSource rd(p.t_file, (int)p.threads, 1 << 20);
            if (!rd.ok()) {
                fprintf(stderr, "Error: failed to build index from %s\n\nCaused by:\n    %s\n", p.t_file.c_str(),
                        rd.error().c_str());
                return 1;
            }
            Batch b;
            while (rd.next(b)) {
                ref.push_back(std::move(b));
                b = Batch();
            }
            if (!rd.error().empty()) {
                fprintf(stderr, "Error: failed to build index from %s\n\nCaused by:\n    %s\n", p.t_file.c_str(),
                        rd.error().c_str());
                return 1;
            }
        }
        if (shm_sketcher_create(p.k, p.h_frac, &sk)) {
            fprintf(stderr, "ERROR: %s\n", shm_last_error());
            return 1;
        }
        std::vector<std::string> names;
        std::vector<shm_segment> segs;
        for (const Batch &bt : ref)
            for (size_t i = 0; i < bt.size(); i++) names.push_back(bt.name(i));
        size_t at = 0;
        for (const Batch &bt : ref)
            for (size_t i = 0; i < bt.size(); i++, at++)
                segs.push_back({names[at].c_str(), bt.bases.data() + bt.off[i], (int64_t)(bt.off[i + 1] - bt.off[i])});
        if (shm_index_build(sk, segs.data(), (int32_t)segs.size(), p.have_max_matches ? p.max_matches : 0, p.device,
                            &idx)) {
            fprintf(stderr, "Error: failed to build index from %s\n\nCaused by:\n    %s\n", p.t_file.c_str(),
                    shm_last_error());
            return 1;
        }
        for (const shm_segment &sg : segs) nucl += (long long)sg.len;
        nsegs = segs.size();
    }
    {
        int64_t ik, ih, hf, bk, bh;
        shm_index_stats(idx, &ik, &ih, &hf, &bk, &bh);
        fprintf(stderr, " | total nucleotides:     %lld\n | index segments:        %zu\n", nucl, nsegs);
        fprintf(stderr, " | indexed kmers:         %lld\n | indexed hits:          %lld\n", (long long)ik, (long long)ih);
        fprintf(stderr, " | | most frequent kmer:      %lld times.\n | | blacklisted kmers:       %lld\n",
                (long long)hf, (long long)bk);
        index_s = now_s() - t_idx;
        fprintf(stderr, " | indexing time:        %.1fs\n", index_s);
    }
    if (!p.save_index.empty()) {
        if (shm_index_save(idx, p.save_index.c_str())) {
            fprintf(stderr, "Error: failed to write the index to %s\n\nCaused by:\n    %s\n", p.save_index.c_str(),
                    shm_last_error());
            return 1;
        }
        fprintf(stderr, "Index written to %s\n", p.save_index.c_str());
    }

    // ---- mapping (src/shmap/mod.rs:374-650)
    shm_map_params mpp{p.k, p.h_frac, p.theta, p.min_diff, p.max_overlap, p.metric, (int32_t)p.nbp,
                       (int32_t)p.one_sweep, (int32_t)p.abs_pos};
    shm_mapper *mp = nullptr;
    if (shm_mapper_create(idx, &mpp, &mp)) {
        fprintf(stderr, "ERROR: %s\n", shm_last_error());
        return 1;
    }
    fprintf(stderr, "Mapping reads using SHmap...\n");
    FILE *unmapped = nullptr, *stats = nullptr;
    if (!p.params_file.empty()) {
        const std::string fn = p.params_file + ".unmapped.paf";
        fprintf(stderr, "Unmapped reads to %s\n", fn.c_str());
        unmapped = fopen(fn.c_str(), "w");
    }
    if (!p.per_read_stats.empty()) {
        fprintf(stderr, "Per-read stats to %s\n", p.per_read_stats.c_str());
        stats = fopen(p.per_read_stats.c_str(), "w");
        if (stats)
            fprintf(stats, "query_id\tread_len\tkmers_sketched\tkmers_seeds\tpossible_matches\texamined_matches\t"
                           "seed_matches\tseeded_buckets\trefined_buckets\tfinal_buckets\tmatches_in_mapping\tmapped\t"
                           "mapq\tquery_mapping_s\n");
    }
    ReadSource rd(p.p_file, (int)p.threads, p.batch);
    if (!rd.ok()) {
        fprintf(stderr, "Error: failed while mapping reads from %s\n\nCaused by:\n    %s\n", p.p_file.c_str(),
                rd.error().c_str());
        return 1;
    }
    static char outbuf[1 << 20];
    setvbuf(stdout, outbuf, _IOFBF, sizeof outbuf);
    std::vector<shm_read> reads;
    std::vector<shm_read_result> res;
    unsigned long long idx_read = 0, n_reads = 0, n_mapped = 0;
    double map_s = 0;
    int rc = 0;
    std::string first_error;  // the first read that failed
    // a batch's records are formatted by up to -@ threads, a slice each, and
    // written in order
    const int nfmt = (int)(p.threads < 1 ? 1 : p.threads > 16 ? 16 : p.threads);
    struct Slice {
        std::string paf, unm, st, err;
        unsigned long long mapped = 0;
    };
    std::vector<Slice> slices(nfmt);
    double read_s = 0, fmt_s = 0;
    {
        BatchReader br([&rd](Batch &b) { return rd.next(b); });
        while (Batch *batch = br.next()) {
            const size_t nb = batch->size();
            reads.resize(nb);
            for (size_t i = 0; i < nb; i++)
                reads[i] = {batch->bases.data() + batch->off[i], (int32_t)(batch->off[i + 1] - batch->off[i])};
            res.resize(nb);
            const double t0 = now_s();
            // A read that fails does so alone (its status); every other read is
            // still written, and the run ends with the first failure, as the
            // reference's collector does (src/shmap/mod.rs:597-612, 637-640).
            if (shm_map_batch(mp, reads.data(), (int32_t)nb, res.data()) && first_error.empty()) {
                size_t bad = 0;
                while (bad < nb && res[bad].status == 0) bad++;
                first_error = bad < nb ? "read " + batch->name(bad) + ": " + shm_last_error() : shm_last_error();
            }
            const double dt = now_s() - t0;
            map_s += dt;
            const double per_read = dt / nb;
            auto format = [&](Slice &sl, size_t i0, size_t i1) {
                sl.paf.clear();
                sl.unm.clear();
                sl.st.clear();
                sl.err.clear();
                sl.mapped = 0;
                std::string line(8192, '\0');
                for (size_t i = i0; i < i1; i++) {
                    const shm_read_result &x = res[i];
                    if (x.status != 0) continue;  // a failed read writes nothing
                    const std::string name = batch->name(i);
                    if (stats && (idx_read + i) % p.per_read_stats_sample == 0) {
                        int n = snprintf(&line[0], line.size(), "%s\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%.9f\n",
                                         name.c_str(), x.read_len, x.m, x.S, x.possible_matches, x.total_matches,
                                         x.seed_matches, x.seeded_buckets, x.refined_buckets, x.final_buckets,
                                         x.mapped ? x.intersection : 0, x.mapped, x.mapped ? (x.mapq == 60 ? 60 : 0) : -1,
                                         per_read);
                        if (n >= (int)line.size()) {
                            line.resize(n + 1);
                            n = snprintf(&line[0], line.size(), "%s\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%.9f\n",
                                         name.c_str(), x.read_len, x.m, x.S, x.possible_matches, x.total_matches,
                                         x.seed_matches, x.seeded_buckets, x.refined_buckets, x.final_buckets,
                                         x.mapped ? x.intersection : 0, x.mapped, x.mapped ? (x.mapq == 60 ? 60 : 0) : -1,
                                         per_read);
                        }
                        sl.st.append(line.data(), (size_t)n);
                    }
                    if (x.mapped) {
                        int n = shm_format_paf(idx, &mpp, name.c_str(), &x, per_read, &line[0], line.size());
                        if (n >= (int)line.size()) {
                            line.resize(n + 1);
                            n = shm_format_paf(idx, &mpp, name.c_str(), &x, per_read, &line[0], line.size());
                        }
                        if (n < 0) {  // not a well-formed mapped record: this read's failure
                            if (sl.err.empty()) sl.err = "read " + name + ": " + shm_last_error();
                            continue;
                        }
                        sl.paf.append(line.data(), (size_t)n);
                        sl.paf.push_back('\n');
                        sl.mapped++;
                    } else if (unmapped) {
                        int n = shm_format_unmapped(name.c_str(), x.read_len, &line[0], line.size());
                        if (n >= (int)line.size()) {
                            line.resize(n + 1);
                            n = shm_format_unmapped(name.c_str(), x.read_len, &line[0], line.size());
                        }
                        sl.unm.append(line.data(), (size_t)n);
                        sl.unm.push_back('\n');
                    }
                }
            };
            const int ns = nb >= 4096 ? nfmt : 1;
            if (ns == 1) {
                format(slices[0], 0, nb);
            } else {
                std::vector<std::thread> th;
                for (int t = 0; t < ns; t++)
                    th.emplace_back([&, t] { format(slices[t], nb * t / ns, nb * (t + 1) / ns); });
                for (std::thread &x : th) x.join();
            }
            for (int t = 0; t < ns; t++) {
                const Slice &sl = slices[t];
                fwrite(sl.paf.data(), 1, sl.paf.size(), stdout);
                if (unmapped) fwrite(sl.unm.data(), 1, sl.unm.size(), unmapped);
                if (stats) fwrite(sl.st.data(), 1, sl.st.size(), stats);
                if (first_error.empty() && !sl.err.empty()) first_error = sl.err;
                n_mapped += sl.mapped;
            }
            idx_read += nb;
            n_reads += nb;
            fmt_s += now_s() - t0 - dt;
            br.recycle(batch);
        }
        read_s = br.busy_s();
    }
    if (!first_error.empty()) {
        fprintf(stderr, "Error: failed while mapping reads from %s\n\nCaused by:\n    %s\n", p.p_file.c_str(),
                first_error.c_str());
        rc = 1;
    }
    if (rc == 0 && !rd.error().empty()) {  // the reader thread has ended with the BatchReader
        fprintf(stderr, "Error: failed while mapping reads from %s\n\nCaused by:\n    %s\n", p.p_file.c_str(),
                rd.error().c_str());
        rc = 1;
    }
    fflush(stdout);
    if (unmapped) fclose(unmapped);
    if (stats) fclose(stats);
    fprintf(stderr, " | reads:                 %llu\n | mapped reads:          %llu\n | mapping time:         %.2fs\n",
            n_reads, n_mapped, map_s);
    if (p.profile) {
        const std::string path = !p.profile_log.empty() ? p.profile_log
                                 : !p.params_file.empty() ? p.params_file + ".profile.json"
                                                          : "shmap.profile.json";
        if (FILE *f = fopen(path.c_str(), "w")) {
            struct stat st;
            const long long in_bytes = stat(p.p_file.c_str(), &st) == 0 ? (long long)st.st_size : -1;
            fprintf(f,
                    "{\"reads\": %llu, \"mapped_reads\": %llu, \"index_s\": %.6f, \"read_s\": %.6f, "
                    "\"input_bytes\": %lld, \"mapping_s\": %.6f, \"output_s\": %.6f, \"total_s\": %.6f, "
                    "\"threads\": %lld}\n",
                    n_reads, n_mapped, index_s, read_s, in_bytes, map_s, fmt_s, now_s() - t_start, p.threads);
            fclose(f);
        }
    }
    shm_mapper_free(mp);
    shm_index_free(idx);
    shm_sketcher_free(sk);
    return rc;
}
