// mps.cpp -- MPS ingestion: the reference's read_mps (src/mps_reader.cpp:
// 54-308) followed by to_ranged_lp (include/ranged_lp.hpp:27-62), and an MPS
// writer for any supr_lp.
//
// The reader gives the reference's RangedLP for every file the reference
// accepts, and fails where it fails:
//  * free format first; if that parse fails anywhere, the file is parsed
//    again as fixed format (fields cut at columns 2-3, 5-12, 15-22, 25-36,
//    40-47, 50-61); if that fails too, the error is the free parse's;
//  * sections NAME, OBJSENSE, ROWS, COLUMNS (INTORG / INTEND markers), RHS
//    (an RHS on the objective row is minus the objective constant), RANGES,
//    BOUNDS (UP LO FX FR MI PL BV LI UI, optional bound-set name),
//    QUADOBJ / QMATRIX / QSECTION (lower triangle kept), ENDATA;
//  * the first N row is the objective, later N rows are free and dropped;
//  * numbers convert as std::stod does (strtod, trailing text ignored, a
//    range error for results that overflow or underflow);
//  * explicit zero coefficients and entries of undeclared rows are dropped;
//    columns are numbered by first appearance; a row or name declared twice
//    keeps the reference's map semantics (the last declaration wins).
// It is one pass over the file held in memory (read with one call) with
// names as views into it and one hash lookup per name, instead of a stream
// and a string map per token; the COLUMNS body of a free-format file, most
// of a large file, is split over threads (ParallelColumns), and the output
// arrays are filled on threads too.  A line of only vertical tabs or form
// feeds in COLUMNS is an error here ("empty COLUMNS line"); the reference
// indexes its missing first token there (src/mps_reader.cpp:175).  Numbers with at most 19 significant digits and a decimal exponent
// within +-22 convert by one exact integer-to-double conversion and one
// correctly rounded multiply or divide by an exact power of ten, which is the
// correctly rounded value strtod also returns; anything else goes to strtod.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <thread>
#include <unordered_map>
#include <vector>

#include <sys/stat.h>

#include "host_lp.hpp"

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

bool is_ws(char ch) {
    return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\v' || ch == '\f' || ch == '\r';
}

// std::stod: leading white space skipped, the longest valid prefix
// converted, std::invalid_argument if none, std::out_of_range on ERANGE.
double parse_double_slow(std::string_view t) {
    char small[64];
    std::string big;
    const char* s;
    if (t.size() < sizeof(small)) {
        std::memcpy(small, t.data(), t.size());
        small[t.size()] = '\0';
        s = small;
    } else {
        big.assign(t.data(), t.size());
        s = big.c_str();
    }
    errno = 0;
    char* end = nullptr;
    const double v = std::strtod(s, &end);
    if (end == s) throw std::invalid_argument("stod");
    if (errno == ERANGE) throw std::out_of_range("stod");
    return v;
}

double parse_double(std::string_view t) {
    static const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                      1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                      1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    const char* p = t.data();
    const char* e = p + t.size();
    bool neg = false;
    if (p < e && *p == '-') { neg = true; ++p; }
    uint64_t mant = 0;
    int digits = 0, frac = 0;
    bool any = false;
    while (p < e && *p >= '0' && *p <= '9') {
        if (mant != 0 || *p != '0') ++digits;
        mant = mant * 10 + (uint64_t)(*p - '0');
        ++p; any = true;
        if (digits > 19) return parse_double_slow(t);
    }
    if (p < e && *p == '.') {
        ++p;
        while (p < e && *p >= '0' && *p <= '9') {
            if (mant != 0 || *p != '0') ++digits;
            mant = mant * 10 + (uint64_t)(*p - '0');
            ++frac; ++p; any = true;
            if (digits > 19) return parse_double_slow(t);
        }
    }
    if (!any) return parse_double_slow(t);
    int ex = 0;
    if (p < e && (*p == 'e' || *p == 'E')) {
        const char* q = p + 1;
        bool eneg = false;
        if (q < e && (*q == '+' || *q == '-')) { eneg = *q == '-'; ++q; }
        if (!(q < e && *q >= '0' && *q <= '9')) return parse_double_slow(t);
        while (q < e && *q >= '0' && *q <= '9') {
            ex = ex * 10 + (*q - '0');
            ++q;
            if (ex > 10000) return parse_double_slow(t);
        }
        if (eneg) ex = -ex;
        p = q;
    }
    if (p != e) return parse_double_slow(t);   // trailing text: strtod decides
    const int e10 = ex - frac;
    if (mant > (uint64_t(1) << 53) || e10 < -22 || e10 > 22) return parse_double_slow(t);
    double v = (double)mant;
    if (e10 < 0) v /= kPow10[-e10];
    else if (e10 > 0) v *= kPow10[e10];
    return neg ? -v : v;
}

// strtod consumed the whole token (the reference's is_number).
bool is_number(std::string_view t) {
    if (t.empty()) return false;
    const std::string s(t);
    char* end = nullptr;
    std::strtod(s.c_str(), &end);
    return end && *end == '\0';
}

std::string_view trim(std::string_view s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t')) ++a;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t')) --b;
    return s.substr(a, b - a);
}

// The line at p (p moves past its '\n'), a trailing '\r' removed.
std::string_view take_line(const char*& p, const char* end) {
    const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
    const char* le = nl ? nl : end;
    std::string_view line(p, (size_t)(le - p));
    p = nl ? nl + 1 : end;
    if (!line.empty() && line.back() == '\r') line.remove_suffix(1);
    return line;
}

// Blank (spaces and tabs only) or a comment ('*' as the first other character).
bool skip_line(std::string_view line) {
    size_t first = 0;
    while (first < line.size() && (line[first] == ' ' || line[first] == '\t')) ++first;
    return first == line.size() || line[first] == '*';
}

void split(std::string_view line, std::vector<std::string_view>& toks) {
    toks.clear();
    for (size_t i = 0; i < line.size();) {
        while (i < line.size() && is_ws(line[i])) ++i;
        if (i == line.size()) break;
        size_t j = i;
        while (j < line.size() && !is_ws(line[j])) ++j;
        toks.push_back(line.substr(i, j - i));
        i = j;
    }
}

constexpr int kEndata = 7;
const char* const kSections[] = {"NAME", "OBJSENSE", "ROWS", "COLUMNS", "RHS", "RANGES",
                                 "BOUNDS", "ENDATA", "QUADOBJ", "QMATRIX", "QSECTION"};

// A section line: a keyword as the first token, in column 1 (ENDATA anywhere).
// Returns the keyword's index in kSections, or -1 for a data line.
int section_keyword(std::string_view line, const std::vector<std::string_view>& toks) {
    if (toks.empty()) return -1;
    int kw = -1;
    for (int k = 0; k < 11; ++k)
        if (toks[0] == kSections[k]) { kw = k; break; }
    const bool col1 = line[0] != ' ' && line[0] != '\t';
    return kw >= 0 && (col1 || kw == kEndata) ? kw : -1;
}

bool unquoted_eq(std::string_view t, const char* word) {
    size_t k = 0;
    const size_t wl = std::strlen(word);
    for (char ch : t) {
        if (ch == '\'') continue;
        if (k >= wl || ch != word[k]) return false;
        ++k;
    }
    return k == wl;
}

// Runs f(0) .. f(T-1) on T threads (f(0) on the caller's); false if any threw.
template <class F>
bool parallel_for(int T, F&& f) {
    std::atomic<bool> ok{true};
    auto body = [&](int t) {
        try { f(t); } catch (...) { ok.store(false); }
    };
    std::vector<std::thread> th;
    try {
        th.reserve((size_t)T - 1);
        for (int t = 1; t < T; ++t) th.emplace_back(body, t);
    } catch (...) {
        ok.store(false);
    }
    body(0);
    for (auto& x : th) x.join();
    return ok.load();
}

// Threads for `work` items, each thread taking at least `unit` of them: the
// hardware's threads (at most 16), or SUPR_MPS_THREADS; SUPR_MPS_MIN_CHUNK
// replaces unit (small values put small files on the parallel paths).
int worker_count(size_t work, size_t unit) {
    long T = (long)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (const char* e = std::getenv("SUPR_MPS_THREADS")) T = std::max(1L, std::strtol(e, nullptr, 10));
    if (const char* e = std::getenv("SUPR_MPS_MIN_CHUNK")) unit = (size_t)std::max(1L, std::strtol(e, nullptr, 10));
    return (int)std::max(1L, std::min(T, (long)(work / unit)));
}

// f(t, lo, hi) for T contiguous ranges of [0, n): on T threads, or one after
// another on the caller's thread when T < 2 or a thread fails (so an
// exception surfaces here; f must give the same result when run again).
template <class F>
void for_ranges(size_t n, int T, F&& f) {
    auto lo = [&](int t) { return n * (size_t)t / (size_t)T; };
    if (T >= 2 && parallel_for(T, [&](int t) { f(t, lo(t), lo(t + 1)); })) return;
    for (int t = 0; t < T; ++t) f(t, lo(t), lo(t + 1));
}

// Names interned to dense ids (views into the file buffer or static text):
// an open-addressing table of (hash, id) slots, linear probing, at most half
// full -- no allocation per name, one cache line per probe.
struct Names {
    struct Slot { uint64_t hash; int64_t id; };   // id -1: empty
    std::vector<std::string_view> name;
    std::vector<Slot> slots;
    size_t mask = 0;

    static uint64_t hash(std::string_view s) {
        uint64_t h = 0x9E3779B97F4A7C15ull ^ (uint64_t)s.size();
        size_t i = 0;
        for (; i + 8 <= s.size(); i += 8) {
            uint64_t w;
            std::memcpy(&w, s.data() + i, 8);
            h = (h ^ w) * 0xFF51AFD7ED558CCDull;
            h ^= h >> 32;
        }
        uint64_t w = 0;
        std::memcpy(&w, s.data() + i, s.size() - i);
        h = (h ^ w) * 0xC4CEB9FE1A85EC53ull;
        return h ^ (h >> 29);
    }
    void grow() {
        const size_t cap = mask ? 2 * (mask + 1) : 1024;
        std::vector<Slot> ns(cap, Slot{0, -1});
        for (const Slot& e : slots) {
            if (e.id < 0) continue;
            size_t q = e.hash & (cap - 1);
            while (ns[q].id >= 0) q = (q + 1) & (cap - 1);
            ns[q] = e;
        }
        slots.swap(ns);
        mask = cap - 1;
    }
    int32_t get(std::string_view s) {
        if (2 * (name.size() + 1) > mask + 1) grow();
        const uint64_t h = hash(s);
        size_t q = h & mask;
        while (slots[q].id >= 0) {
            if (slots[q].hash == h && name[slots[q].id] == s) return (int32_t)slots[q].id;
            q = (q + 1) & mask;
        }
        const int32_t k = (int32_t)name.size();
        slots[q] = Slot{h, k};
        name.push_back(s);
        return k;
    }
    int32_t find(std::string_view s) const {
        if (!mask) return -1;
        const uint64_t h = hash(s);
        size_t q = h & mask;
        while (slots[q].id >= 0) {
            if (slots[q].hash == h && name[slots[q].id] == s) return (int32_t)slots[q].id;
            q = (q + 1) & mask;
        }
        return -1;
    }
    // The table of nm (distinct names, id = position; h their hashes), filled
    // on T threads: each claims a free slot by compare-and-swap of its id and
    // then writes the hash.  Slot order depends on the threads' timing, ids
    // and lookups do not.  False if a thread could not run.
    bool build(std::vector<std::string_view>&& nm, const std::vector<uint64_t>& h, int T) {
        name = std::move(nm);
        size_t cap = 1024;
        while (2 * (name.size() + 1) > cap) cap *= 2;
        slots.assign(cap, Slot{0, -1});
        mask = cap - 1;
        const size_t n = name.size();
        return parallel_for(T, [&](int t) {
                const size_t g1 = n * (size_t)(t + 1) / (size_t)T;
                for (size_t g = n * (size_t)t / (size_t)T; g < g1; ++g) {
                    if (g + 16 < g1) __builtin_prefetch(&slots[h[g + 16] & mask], 1);
                    size_t q = h[g] & mask;
                    for (;; q = (q + 1) & mask) {
                        int64_t none = -1;
                        if (__atomic_compare_exchange_n(&slots[q].id, &none, (int64_t)g, false,
                                                        __ATOMIC_RELAXED, __ATOMIC_RELAXED))
                            break;
                    }
                    slots[q].hash = h[g];
                }
            });
    }
};

// One coefficient of A as read: row-name id, column id, value.
struct Entry { int32_t row, col; double val; };

// The body of a free-format COLUMNS section parsed on several threads, with
// the result of the one-thread parse: the same column numbering (first
// appearance), the same entries in file order, the objective summed in file
// order.  The body is cut at line boundaries into one chunk per thread; each
// thread reads its chunk as COLUMNS lines until a section line, splitting
// it into runs (consecutive lines of one column name), and gives up at
// anything the one-thread parse must see in order: an INTORG / INTEND
// marker, a row name not yet declared, a number strtod rejects, an empty
// line.  The first chunk that meets a section line ends the body (later
// chunks read past it and are dropped); if a chunk gives up before that,
// run() returns false, nothing is kept, and the caller parses the body
// itself, so every error and every unusual file takes the one-thread path.
// Runs are then numbered across chunks through hash-partitioned tables (a
// name's first run gets the next id, later runs of it share that id) and the
// column table is built from the ids on all threads.
struct ParallelColumns {
    Names cols;
    std::vector<double> obj;
    std::vector<Entry> entries;
    const char* resume = nullptr;

    struct Chunk {
        const char* begin = nullptr;
        const char* end = nullptr;
        bool stopped = false, gave_up = false;
        const char* stop = nullptr;            // the section line that ended the body
        std::vector<std::string_view> run_name;
        std::vector<uint64_t> run_hash;
        struct Coef { int32_t row, run; double val; };
        std::vector<Coef> coef;
        struct Cost { int32_t run; double val; };
        std::vector<Cost> cost;
        std::vector<std::vector<int32_t>> part;   // per name partition: this chunk's runs in it
        std::vector<int64_t> first;            // per run: -1, or (chunk << 32 | run) of the name's first run
        std::vector<int32_t> id;               // per run: column id
        int32_t fresh = 0;                     // runs here that are a name's first
    };

    // The partition of a name hash (the table inside a partition uses the low bits).
    static int part_of(uint64_t h, int Q) { return (int)((h >> 40) % (uint64_t)Q); }

    static void scan(Chunk& C, int index, std::atomic<int>& first_end, const Names& rows,
                     const std::vector<char>& rfree, bool has_obj, std::string_view obj_name) {
        auto finish = [&](bool stopped) {
            (stopped ? C.stopped : C.gave_up) = true;
            int cur = first_end.load();
            while (index < cur && !first_end.compare_exchange_weak(cur, index)) {}
        };
        std::vector<std::string_view> toks;
        toks.reserve(16);
        std::string_view run;
        const char* p = C.begin;
        size_t lines = 0;
        try {
            while (p < C.end) {
                // a chunk before this one has ended the body: this one is dropped
                if ((++lines & 1023) == 0 && first_end.load(std::memory_order_relaxed) < index) return;
                const char* at = p;
                const std::string_view line = take_line(p, C.end);
                if (skip_line(line)) continue;
                split(line, toks);
                if (section_keyword(line, toks) >= 0) { C.stop = at; finish(true); return; }
                if (toks.empty()) { finish(false); return; }
                for (auto& x : toks)
                    if (unquoted_eq(x, "MARKER")) { finish(false); return; }
                if (C.run_name.empty() || toks[0] != run) {
                    run = toks[0];
                    const uint64_t h = Names::hash(run);
                    C.part[part_of(h, (int)C.part.size())].push_back((int32_t)C.run_name.size());
                    C.run_name.push_back(run);
                    C.run_hash.push_back(h);
                }
                const int32_t r = (int32_t)C.run_name.size() - 1;
                for (size_t i = 1; i + 1 < toks.size(); i += 2) {
                    const double v = parse_double(toks[i + 1]);
                    if (has_obj && toks[i] == obj_name) { C.cost.push_back({r, v}); continue; }
                    const int32_t k = rows.find(toks[i]);
                    if (k < 0) { finish(false); return; }
                    if (rfree[k]) continue;
                    C.coef.push_back({k, r, v});
                }
            }
        } catch (...) {
            finish(false);
        }
    }

    // p: the line after the COLUMNS line.  False: nothing parsed, the caller
    // reads the body itself.
    bool run(const char* p, const char* end, const Names& rows, const std::vector<char>& rfree,
             bool has_obj, std::string_view obj_name) {
        const int T = worker_count((size_t)(end - p), 1 << 20);   // at least 1 MiB of body each
        if (T < 2) return false;
        std::vector<Chunk> ch((size_t)T);
        const char* at = p;
        for (int t = 0; t < T; ++t) {
            ch[t].begin = at;
            const char* cut = t + 1 == T ? end : p + (size_t)(end - p) * (size_t)(t + 1) / (size_t)T;
            if (cut < at) cut = at;
            if (cut < end) {
                const char* nl = static_cast<const char*>(std::memchr(cut, '\n', (size_t)(end - cut)));
                cut = nl ? nl + 1 : end;
            }
            ch[t].end = at = cut;
        }
        for (auto& c : ch) c.part.resize((size_t)T);
        std::atomic<int> first_end{INT_MAX};
        if (!parallel_for(T, [&](int t) { scan(ch[t], t, first_end, rows, rfree, has_obj, obj_name); }))
            return false;
        // chunks 0 .. last-1 are whole COLUMNS lines; last ends the body or gave up
        int last = first_end.load();
        if (last == INT_MAX) last = T - 1;
        if (ch[last].gave_up) return false;
        resume = ch[last].stopped ? ch[last].stop : end;
        const int nc = last + 1;

        // number the runs: partition q owns the names whose hash's top bits say q
        size_t runs = 0;
        for (int c = 0; c < nc; ++c) {
            ch[c].first.assign(ch[c].run_name.size(), -1);
            runs += ch[c].run_name.size();
        }
        if (!parallel_for(T, [&](int q) {
                size_t cnt = 0;
                for (int c = 0; c < nc; ++c) cnt += ch[c].part[q].size();
                size_t cap = 64;
                while (cap < 2 * cnt + 2) cap *= 2;
                struct Seen { uint64_t hash; int64_t at; };   // at -1: empty
                std::vector<Seen> tab(cap, Seen{0, -1});
                for (int c = 0; c < nc; ++c) {
                    const Chunk& C = ch[c];
                    for (const int32_t r : C.part[q]) {
                        const uint64_t h = C.run_hash[r];
                        size_t s = h & (cap - 1);
                        for (;; s = (s + 1) & (cap - 1)) {
                            if (tab[s].at < 0) { tab[s] = Seen{h, ((int64_t)c << 32) | (int64_t)r}; break; }
                            const int64_t a = tab[s].at;
                            if (tab[s].hash == h && ch[a >> 32].run_name[a & 0xffffffff] == C.run_name[r]) {
                                ch[c].first[r] = a;
                                break;
                            }
                        }
                    }
                }
            }))
            return false;
        std::vector<int32_t> base((size_t)nc + 1, 0);
        for (int c = 0; c < nc; ++c) {
            int32_t f = 0;
            for (int64_t a : ch[c].first) f += a < 0;
            ch[c].fresh = f;
            base[c + 1] = base[c] + f;
        }
        const int32_t n = base[nc];
        std::vector<std::string_view> name((size_t)n);
        std::vector<uint64_t> hash((size_t)n);
        if (!parallel_for(nc, [&](int c) {
                Chunk& C = ch[c];
                C.id.resize(C.first.size());
                int32_t g = base[c];
                for (size_t r = 0; r < C.first.size(); ++r)
                    if (C.first[r] < 0) {
                        C.id[r] = g;
                        name[g] = C.run_name[r];
                        hash[g] = C.run_hash[r];
                        ++g;
                    }
            }))
            return false;
        if (!parallel_for(nc, [&](int c) {
                Chunk& C = ch[c];
                for (size_t r = 0; r < C.first.size(); ++r)
                    if (C.first[r] >= 0) C.id[r] = ch[C.first[r] >> 32].id[C.first[r] & 0xffffffff];
            }))
            return false;
        if (!cols.build(std::move(name), hash, T)) return false;

        obj.assign((size_t)n, 0.0);
        for (int c = 0; c < nc; ++c)   // file order, as the one-thread parse adds
            for (const auto& e : ch[c].cost) obj[ch[c].id[e.run]] += e.val;
        std::vector<size_t> off((size_t)nc + 1, 0);
        for (int c = 0; c < nc; ++c) off[c + 1] = off[c] + ch[c].coef.size();
        entries.resize(off[nc]);
        return parallel_for(nc, [&](int c) {
            const Chunk& C = ch[c];
            Entry* out = entries.data() + off[c];
            for (size_t i = 0; i < C.coef.size(); ++i)
                out[i] = Entry{C.coef[i].row, C.id[C.coef[i].run], C.coef[i].val};
        });
    }
};

struct Parsed {
    std::string name;
    bool maximize = false;
    std::vector<std::string> col_names, row_names;
    std::string obj_name;
    std::vector<int32_t> row, col;
    std::vector<double> val, c, rL, rU, l, u;
    double obj_offset = 0.0;
    std::vector<char> integer;
    std::vector<int32_t> q_row, q_col;
    std::vector<double> q_val;
};

enum Section { S_NONE, S_OBJSENSE, S_ROWS, S_COLUMNS, S_RHS, S_RANGES, S_BOUNDS, S_QUAD, S_QMATRIX };

[[noreturn]] void fail(const std::string& msg) { throw std::runtime_error(msg); }

// One parse of the whole text (src/mps_reader.cpp:65-307 + to_ranged_lp).
void parse(std::string_view text, bool fixed_format, Parsed& out) {
    std::string sense = "min";
    std::string_view pname;
    bool have_name = false;
    Section section = S_NONE;

    Names rows;                       // every row name seen (ROWS, COLUMNS, RHS, RANGES)
    std::vector<char> rtype;          // per row-name id: type char (0 = undeclared)
    std::vector<char> rdecl;          // per row-name id: declared in ROWS
    std::vector<char> rfree;          // per row-name id: a free (second or later N) row
    std::vector<int32_t> row_order;   // ROWS order of non-N rows (row-name ids)
    int32_t obj_id = -1;
    std::string_view obj_name;

    Names cols;
    std::vector<double> obj;          // per column: objective coefficient
    std::vector<char> integ, has_lo, has_hi, explicit_lo;
    std::vector<double> blo, bhi;
    bool int_marker = false;
    int32_t last_col = -1;            // the column of the previous COLUMNS line
    std::string_view last_col_name;
    auto add_col = [&](std::string_view nm) {
        const int32_t before = (int32_t)cols.name.size();
        const int32_t j = cols.get(nm);
        if (j == before) {
            obj.push_back(0.0); integ.push_back(0); has_lo.push_back(0); has_hi.push_back(0);
            explicit_lo.push_back(0); blo.push_back(0.0); bhi.push_back(0.0);
        }
        return j;
    };
    auto row_id = [&](std::string_view nm) {
        const int32_t before = (int32_t)rows.name.size();
        const int32_t k = rows.get(nm);
        if (k == before) { rtype.push_back(0); rdecl.push_back(0); rfree.push_back(0); }
        return k;
    };

    std::vector<Entry> entries;
    std::vector<double> rhs_v, rng_v;   // per row-name id
    std::vector<char> rhs_has, rng_has;
    double obj_constant = 0.0;
    struct QEntry { std::string_view a, b; double v; bool full; };
    std::vector<QEntry> quad;

    std::vector<std::string_view> toks, tk;
    toks.reserve(16);
    tk.reserve(16);
    const char* p = text.data();
    const char* const end = p + text.size();
    while (p < end) {
        const std::string_view line = take_line(p, end);
        if (skip_line(line)) continue;
        split(line, toks);
        const int kw = section_keyword(line, toks);
        if (kw >= 0) {
            if (kw == 0) {   // NAME
                have_name = true;
                pname = toks.size() > 1 ? toks[1] : std::string_view("UNNAMED");
                section = S_NONE;
                continue;
            }
            if (kw == kEndata) break;
            static const Section map[] = {S_NONE, S_OBJSENSE, S_ROWS, S_COLUMNS, S_RHS, S_RANGES,
                                          S_BOUNDS, S_NONE, S_QUAD, S_QMATRIX, S_QUAD};
            section = map[kw];
            // the first COLUMNS section of a free-format file, before any
            // column is known: its body is parsed on several threads
            if (section == S_COLUMNS && !fixed_format && cols.name.empty() && !int_marker) {
                ParallelColumns pc;
                if (pc.run(p, end, rows, rfree, obj_id >= 0, obj_name)) {
                    cols = std::move(pc.cols);
                    const size_t n = cols.name.size();
                    obj = std::move(pc.obj);
                    integ.assign(n, 0); has_lo.assign(n, 0); has_hi.assign(n, 0);
                    explicit_lo.assign(n, 0); blo.assign(n, 0.0); bhi.assign(n, 0.0);
                    entries = std::move(pc.entries);
                    p = pc.resume;   // the section line that ended the body, or the end
                }
            }
            continue;
        }
        const std::vector<std::string_view>* tp = &toks;
        if (fixed_format && section != S_OBJSENSE) {
            static const int st[6] = {1, 4, 14, 24, 39, 49};
            static const int ln[6] = {2, 8, 8, 12, 8, 12};
            std::string_view f[6];
            for (int k = 0; k < 6; ++k) {
                if ((int)line.size() <= st[k]) break;
                f[k] = trim(line.substr((size_t)st[k], (size_t)ln[k]));
            }
            tk.clear();
            if (section == S_ROWS) {
                tk = {f[0], f[1]};
            } else if (section == S_BOUNDS) {
                tk = {f[0], f[1].empty() ? std::string_view("BND") : f[1], f[2]};
                if (!f[3].empty()) tk.push_back(f[3]);
            } else {
                std::string_view lead = f[1];
                if (lead.empty() && section != S_COLUMNS) lead = "RHS";
                tk = {lead, f[2], f[3]};
                if (!f[4].empty()) { tk.push_back(f[4]); tk.push_back(f[5]); }
            }
            tp = &tk;
        }
        const std::vector<std::string_view>& t = *tp;
        switch (section) {
        case S_OBJSENSE: {
            if (t.empty()) fail("empty OBJSENSE line");
            std::string s(t[0]);
            for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
            sense = s;
            break;
        }
        case S_ROWS: {
            if (!fixed_format && t.size() != 2) fail("ROWS line is not free format: " + std::string(line));
            const char ty = t[0].empty() ? '\0' : (char)std::toupper((unsigned char)t[0][0]);
            const int32_t k = row_id(t[1]);
            rtype[k] = ty;
            rdecl[k] = 1;
            if (ty == 'N') {
                if (obj_id < 0) { obj_id = k; obj_name = t[1]; }
                else rfree[k] = 1;
            } else {
                row_order.push_back(k);
            }
            break;
        }
        case S_COLUMNS: {
            bool marker = false;
            for (auto& x : toks) if (unquoted_eq(x, "MARKER")) marker = true;
            if (marker) {
                for (auto& x : toks) {
                    if (unquoted_eq(x, "INTORG")) int_marker = true;
                    if (unquoted_eq(x, "INTEND")) int_marker = false;
                }
                break;
            }
            if (t.empty()) fail("empty COLUMNS line");
            // a column's lines are usually consecutive: reuse the last lookup
            if (last_col < 0 || t[0] != last_col_name) { last_col = add_col(t[0]); last_col_name = t[0]; }
            const int32_t j = last_col;
            if (int_marker) integ[j] = 1;
            for (size_t i = 1; i + 1 < t.size(); i += 2) {
                const double v = parse_double(t[i + 1]);
                if (obj_id >= 0 && t[i] == obj_name) { obj[j] += v; continue; }
                const int32_t k = row_id(t[i]);
                if (rfree[k]) continue;
                entries.push_back({k, j, v});
            }
            break;
        }
        case S_RHS:
        case S_RANGES: {
            const size_t s0 = (t.size() % 2 == 1) ? 1 : 0;
            for (size_t i = s0; i + 1 < t.size(); i += 2) {
                const double v = parse_double(t[i + 1]);
                if (section == S_RHS && t[i] == obj_name && obj_id >= 0) { obj_constant = -v; continue; }
                const int32_t k = row_id(t[i]);
                if (section == S_RHS) {
                    if ((size_t)k >= rhs_v.size()) { rhs_v.resize(k + 1, 0.0); rhs_has.resize(k + 1, 0); }
                    rhs_v[k] = v; rhs_has[k] = 1;
                } else {
                    if ((size_t)k >= rng_v.size()) { rng_v.resize(k + 1, 0.0); rng_has.resize(k + 1, 0); }
                    rng_v[k] = v; rng_has[k] = 1;
                }
            }
            break;
        }
        case S_QUAD:
        case S_QMATRIX:
            if (t.size() < 3) break;
            quad.push_back({t[0], t[1], parse_double(t[2]), section == S_QMATRIX});
            break;
        case S_BOUNDS: {
            if (t.empty()) fail("empty BOUNDS line");
            std::string bt(t[0]);
            for (auto& ch : bt) ch = (char)std::toupper((unsigned char)ch);
            const bool needs_val = bt == "UP" || bt == "LO" || bt == "FX" || bt == "LI" || bt == "UI";
            bool has_set;
            if (needs_val) has_set = t.size() >= 4;
            else if (bt == "BV")
                has_set = t.size() >= 4 || (t.size() == 3 && !(cols.find(t[1]) >= 0 && is_number(t[2])));
            else has_set = t.size() >= 3;
            const size_t ci = has_set ? 2 : 1;
            if (t.size() <= ci) fail("malformed BOUNDS line: " + std::string(line));
            const int32_t j = add_col(t[ci]);
            const double v = t.size() > ci + 1 ? parse_double(t[ci + 1]) : 0.0;
            if (bt == "UP") { bhi[j] = v; has_hi[j] = 1; }
            else if (bt == "LI") { blo[j] = v; has_lo[j] = 1; explicit_lo[j] = 1; integ[j] = 1; }
            else if (bt == "UI") { bhi[j] = v; has_hi[j] = 1; integ[j] = 1; }
            else if (bt == "LO") { blo[j] = v; has_lo[j] = 1; explicit_lo[j] = 1; }
            else if (bt == "FX") { blo[j] = v; bhi[j] = v; has_lo[j] = has_hi[j] = 1; explicit_lo[j] = 1; }
            else if (bt == "FR") { blo[j] = -kInf; bhi[j] = kInf; has_lo[j] = has_hi[j] = 1; explicit_lo[j] = 1; }
            else if (bt == "MI") { blo[j] = -kInf; has_lo[j] = 1; explicit_lo[j] = 1; }
            else if (bt == "PL") { bhi[j] = kInf; has_hi[j] = 1; }
            else if (bt == "BV") { blo[j] = 0.0; bhi[j] = 1.0; has_lo[j] = has_hi[j] = 1; explicit_lo[j] = 1; integ[j] = 1; }
            else fail("Unsupported BOUNDS type '" + bt + "' in MPS file");
            break;
        }
        case S_NONE:
            break;
        }
    }
    if (obj_id < 0) fail("No objective (N) row found in MPS file");

    // rows: ROWS order; a name listed twice maps to its last position
    const int32_t m = (int32_t)row_order.size(), n = (int32_t)cols.name.size();
    std::vector<int32_t> pos(rows.name.size(), -1);
    for (int32_t i = 0; i < m; ++i) pos[row_order[i]] = i;
    // explicit zeros and rows never declared in ROWS are dropped, order kept
    {
        const int T = worker_count(entries.size(), 1 << 17);
        std::vector<size_t> kept((size_t)T + 1, 0);
        for_ranges(entries.size(), T, [&](int t, size_t lo, size_t hi) {
            size_t k = 0;
            for (size_t e = lo; e < hi; ++e) k += entries[e].val != 0.0 && pos[entries[e].row] >= 0;
            kept[t + 1] = k;
        });
        for (int t = 0; t < T; ++t) kept[t + 1] += kept[t];
        out.row.resize(kept[T]); out.col.resize(kept[T]); out.val.resize(kept[T]);
        for_ranges(entries.size(), T, [&](int t, size_t lo, size_t hi) {
            size_t k = kept[t];
            for (size_t e = lo; e < hi; ++e) {
                const Entry& en = entries[e];
                const int32_t i = pos[en.row];
                if (en.val == 0.0 || i < 0) continue;
                out.row[k] = i; out.col[k] = en.col; out.val[k] = en.val;
                ++k;
            }
        });
    }
    // to_ranged_lp (include/ranged_lp.hpp:27-62)
    out.maximize = sense == "max";
    out.c.assign(n, 0.0);
    for (int32_t j = 0; j < n; ++j) out.c[j] = out.maximize ? -obj[j] : obj[j];
    out.l.assign(n, 0.0);
    out.u.assign(n, kInf);
    for (int32_t j = 0; j < n; ++j) {
        if (has_lo[j]) out.l[j] = blo[j];
        if (has_hi[j]) out.u[j] = bhi[j];
        if (has_hi[j] && bhi[j] < 0 && !explicit_lo[j]) out.l[j] = -kInf;
    }
    out.rL.assign(m, 0.0);
    out.rU.assign(m, 0.0);
    for (int32_t i = 0; i < m; ++i) {
        const int32_t k = row_order[i];
        const char ty = rtype[k];
        // a right-hand side goes to its name's last position only; types and
        // ranges apply at every position of a twice-declared name
        const double rhs = pos[k] == i && (size_t)k < rhs_has.size() && rhs_has[k] ? rhs_v[k] : 0.0;
        const bool has_range = (size_t)k < rng_has.size() && rng_has[k];
        double& lo = out.rL[i];
        double& hi = out.rU[i];
        if (!has_range) {
            if (ty == 'L') { lo = -kInf; hi = rhs; }
            else if (ty == 'G') { lo = rhs; hi = kInf; }
            else if (ty == 'E') { lo = rhs; hi = rhs; }
        } else {
            const double rng = rng_v[k], r = std::fabs(rng);
            if (ty == 'L') { lo = rhs - r; hi = rhs; }
            else if (ty == 'G') { lo = rhs; hi = rhs + r; }
            else if (ty == 'E') {
                if (rng >= 0) { lo = rhs; hi = rhs + r; }
                else { lo = rhs - r; hi = rhs; }
            }
        }
    }
    out.obj_offset = out.maximize ? -obj_constant : obj_constant;
    out.integer.assign(integ.begin(), integ.end());
    out.q_row.clear(); out.q_col.clear(); out.q_val.clear();
    for (const QEntry& q : quad) {
        int32_t i = cols.find(q.a), j = cols.find(q.b);
        if (i < 0 || j < 0) fail("QUADOBJ refers to unknown column");
        if (q.full && i < j) continue;
        if (i < j) std::swap(i, j);
        out.q_row.push_back(i);
        out.q_col.push_back(j);
        out.q_val.push_back(out.maximize ? -q.v : q.v);
    }
    out.name = have_name && !pname.empty() ? std::string(pname) : std::string("UNNAMED");
    out.obj_name = std::string(obj_name);
    out.col_names.resize(n);
    for_ranges((size_t)n, worker_count((size_t)n, 1 << 17), [&](int, size_t lo, size_t hi) {
        for (size_t j = lo; j < hi; ++j) out.col_names[j].assign(cols.name[j].data(), cols.name[j].size());
    });
    out.row_names.resize(m);
    for (int32_t i = 0; i < m; ++i) out.row_names[i] = std::string(rows.name[row_order[i]]);
}

// The whole file in memory: one read of its size for a regular file, else
// read in blocks (a pipe).  A read error ends the text where it stopped, as
// the stream the reference reads from ends.
struct FileText {
    std::unique_ptr<char[]> buf;
    size_t size = 0;
    std::string_view view() const { return std::string_view(buf.get(), size); }
};

FileText read_file(const char* path) {
    std::FILE* f = std::fopen(path, "rb");
    if (!f) fail(std::string("cannot open MPS file: ") + path);
    struct stat st;
    size_t cap = 1 << 16;
    if (fstat(fileno(f), &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0)
        cap = (size_t)st.st_size + 1;   // one byte more: the end shows as a short read
    FileText t;
    try {
        t.buf.reset(new char[cap]);
        for (;;) {
            const size_t k = std::fread(t.buf.get() + t.size, 1, cap - t.size, f);
            t.size += k;
            if (t.size < cap) break;   // end of file or a read error
            std::unique_ptr<char[]> nb(new char[2 * cap]);
            std::memcpy(nb.get(), t.buf.get(), t.size);
            t.buf = std::move(nb);
            cap *= 2;
        }
    } catch (...) {
        std::fclose(f);
        throw;
    }
    std::fclose(f);
    return t;
}

}  // namespace

struct supr_mps {
    Parsed p;
};

extern "C" {

int supr_mps_read(const char* path, supr_mps** out) {
    if (!path || !out) {
        supr::set_error("supr_mps_read: null argument");
        return SUPR_EINVAL;
    }
    *out = nullptr;
    try {
        auto h = std::make_unique<supr_mps>();
        std::string free_err;
        try {
            const FileText file = read_file(path);
            const std::string_view text = file.view();
            try {
                parse(text, false, h->p);
            } catch (const std::exception& e) {
                free_err = e.what();
                parse(text, true, h->p);
            }
        } catch (const std::bad_alloc&) {
            throw;
        } catch (const std::exception& e) {
            supr::set_error(std::string("MPS parse failed: ") + (free_err.empty() ? e.what() : free_err));
            return SUPR_EINVAL;
        }
        *out = h.release();
        return SUPR_OK;
    } catch (const std::bad_alloc&) {
        supr::set_error("out of host memory");
        return SUPR_ENOMEM;
    } catch (const std::exception& e) {
        supr::set_error(e.what());
        return SUPR_EINTERNAL;
    }
}

int supr_mps_sizes(const supr_mps* h, int32_t* m, int32_t* n, int32_t* nnz, int32_t* q_nnz,
                   int32_t* maximize) {
    if (!h) {
        supr::set_error("supr_mps_sizes: null handle");
        return SUPR_EINVAL;
    }
    if (m) *m = (int32_t)h->p.row_names.size();
    if (n) *n = (int32_t)h->p.col_names.size();
    if (nnz) *nnz = (int32_t)h->p.val.size();
    if (q_nnz) *q_nnz = (int32_t)h->p.q_val.size();
    if (maximize) *maximize = h->p.maximize ? 1 : 0;
    return SUPR_OK;
}

int supr_mps_get(const supr_mps* h, int32_t* row_idx, int32_t* col_idx, double* val, double* c,
                 double* rL, double* rU, double* l, double* u, double* obj_offset, char* integer,
                 int32_t* q_row, int32_t* q_col, double* q_val) {
    if (!h) {
        supr::set_error("supr_mps_get: null handle");
        return SUPR_EINVAL;
    }
    const Parsed& p = h->p;
    auto put = [](auto* dst, const auto& v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    put(row_idx, p.row); put(col_idx, p.col); put(val, p.val);
    put(c, p.c); put(rL, p.rL); put(rU, p.rU); put(l, p.l); put(u, p.u);
    put(integer, p.integer);
    put(q_row, p.q_row); put(q_col, p.q_col); put(q_val, p.q_val);
    if (obj_offset) *obj_offset = p.obj_offset;
    return SUPR_OK;
}

const char* supr_mps_name(const supr_mps* h, int32_t kind, int32_t index) {
    if (!h) return nullptr;
    const Parsed& p = h->p;
    switch (kind) {
    case 0: return p.name.c_str();
    case 1: return index >= 0 && (size_t)index < p.col_names.size() ? p.col_names[index].c_str() : nullptr;
    case 2: return index >= 0 && (size_t)index < p.row_names.size() ? p.row_names[index].c_str() : nullptr;
    case 3: return p.obj_name.c_str();
    default: return nullptr;
    }
}

int supr_mps_names(const supr_mps* h, int32_t kind, char* buf, int64_t cap, int64_t* bytes) {
    if (!h || !bytes || (kind != 1 && kind != 2)) {
        supr::set_error("supr_mps_names: null handle / bytes, or kind not 1 (columns) or 2 (rows)");
        return SUPR_EINVAL;
    }
    const std::vector<std::string>& v = kind == 1 ? h->p.col_names : h->p.row_names;
    int64_t total = 0;
    for (const std::string& x : v) total += (int64_t)x.size() + 1;
    *bytes = total;
    if (!buf) return SUPR_OK;
    if (cap < total) {
        supr::set_error("supr_mps_names: buffer of " + std::to_string(cap) + " bytes, " +
                        std::to_string(total) + " needed");
        return SUPR_EINVAL;
    }
    char* o = buf;
    for (const std::string& x : v) {
        std::memcpy(o, x.data(), x.size());
        o += x.size();
        *o++ = '\0';
    }
    return SUPR_OK;
}

void supr_mps_free(supr_mps* h) { delete h; }

int supr_mps_write(const char* path, const supr_lp* lp, const char* name,
                   const char* const* col_names, const char* const* row_names,
                   const char* integer) {
    if (!path || !supr::check_lp(lp, "supr_mps_write")) {
        if (!path) supr::set_error("supr_mps_write: null path");
        return SUPR_EINVAL;
    }
    try {
        const int32_t m = lp->m, n = lp->n, nnz = lp->nnz;
        auto bad_name = [](const char* nm) {
            if (!nm || !*nm) return true;
            for (const char* q = nm; *q; ++q) if (is_ws(*q)) return true;
            return false;
        };
        std::vector<std::string> cn(n), rn(m);
        for (int32_t j = 0; j < n; ++j) {
            cn[j] = col_names ? std::string(col_names[j] ? col_names[j] : "") : "C" + std::to_string(j);
            if (bad_name(cn[j].c_str())) throw std::invalid_argument("supr_mps_write: column name empty or with white space");
        }
        std::unordered_map<std::string, int> seen;
        for (int32_t i = 0; i < m; ++i) {
            rn[i] = row_names ? std::string(row_names[i] ? row_names[i] : "") : "R" + std::to_string(i);
            if (bad_name(rn[i].c_str())) throw std::invalid_argument("supr_mps_write: row name empty or with white space");
            seen[rn[i]] = 1;
        }
        std::string obj = "COST";
        for (int k = 0; seen.count(obj); ++k) obj = "OBJ" + std::to_string(k);
        for (int32_t i = 0; i < m; ++i)
            if (!(lp->rL[i] <= lp->rU[i]) || lp->rL[i] == kInf || lp->rU[i] == -kInf)
                throw std::invalid_argument("supr_mps_write: row " + std::to_string(i) + " has no feasible range");
        // entries grouped by column, COO order within a column
        std::vector<int32_t> ptr(n + 1, 0), ord(nnz);
        for (int32_t k = 0; k < nnz; ++k) ++ptr[lp->col_idx[k] + 1];
        for (int32_t j = 0; j < n; ++j) ptr[j + 1] += ptr[j];
        {
            std::vector<int32_t> next(ptr.begin(), ptr.end() - 1);
            for (int32_t k = 0; k < nnz; ++k) ord[next[lp->col_idx[k]]++] = k;
        }
        std::FILE* f = std::fopen(path, "wb");
        if (!f) throw std::runtime_error(std::string("cannot open ") + path + " for writing");
        std::string buf;
        buf.reserve(1 << 20);
        char num[64];
        auto g17 = [&](double v) {
            std::snprintf(num, sizeof(num), "%.17g", v);
            return num;
        };
        auto flush = [&](bool force) {
            if (force || buf.size() > (1u << 20)) {
                if (std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) {
                    std::fclose(f);
                    throw std::runtime_error(std::string("write failed: ") + path);
                }
                buf.clear();
            }
        };
        buf += "NAME "; buf += (name && *name && !bad_name(name)) ? name : "LP"; buf += "\n";
        buf += "ROWS\n N  "; buf += obj; buf += "\n";
        // row kinds: E (rL == rU), L (rL = -inf), G (rU = +inf or a range
        // rL + (rU - rL) reproduces), L with a range (rU - (rU - rL)
        // reproduces rL), free rows as L with an infinite right-hand side
        std::vector<char> kind(m);
        std::vector<double> rhs(m), rng(m);
        std::vector<char> has_rng(m, 0);
        for (int32_t i = 0; i < m; ++i) {
            const double lo = lp->rL[i], hi = lp->rU[i];
            if (lo == hi) { kind[i] = 'E'; rhs[i] = lo; }
            else if (lo == -kInf) { kind[i] = 'L'; rhs[i] = hi; }
            else if (hi == kInf) { kind[i] = 'G'; rhs[i] = lo; }
            else {
                const double r = hi - lo;
                has_rng[i] = 1;
                rng[i] = r;
                if (lo + r == hi || !(hi - r == lo)) { kind[i] = 'G'; rhs[i] = lo; }
                else { kind[i] = 'L'; rhs[i] = hi; }
            }
            buf += ' '; buf += kind[i]; buf += "  "; buf += rn[i]; buf += '\n';
            flush(false);
        }
        buf += "COLUMNS\n";
        bool in_int = false;
        for (int32_t j = 0; j < n; ++j) {
            const bool ij = integer && integer[j];
            if (ij != in_int) {
                buf += ij ? "    MARKER  'MARKER'  'INTORG'\n" : "    MARKER  'MARKER'  'INTEND'\n";
                in_int = ij;
            }
            // two (row, value) pairs per line, the objective first
            buf += "    "; buf += cn[j]; buf += "  "; buf += obj; buf += "  "; buf += g17(lp->c[j]);
            bool open = true;
            for (int32_t q = ptr[j]; q < ptr[j + 1]; ++q) {
                const int32_t k = ord[q];
                if (!open) { buf += "    "; buf += cn[j]; }
                buf += "  "; buf += rn[lp->row_idx[k]]; buf += "  "; buf += g17(lp->val[k]);
                if (open) buf += '\n';
                open = !open;
            }
            if (open) buf += '\n';
            flush(false);
        }
        if (in_int) buf += "    MARKER  'MARKER'  'INTEND'\n";
        buf += "RHS\n";
        if (lp->obj_offset != 0.0) { buf += "    RHS  "; buf += obj; buf += "  "; buf += g17(-lp->obj_offset); buf += '\n'; }
        for (int32_t i = 0; i < m; ++i) {
            if (rhs[i] == 0.0 && !std::signbit(rhs[i])) continue;
            buf += "    RHS  "; buf += rn[i]; buf += "  "; buf += g17(rhs[i]); buf += '\n';
            flush(false);
        }
        bool any_rng = false;
        for (int32_t i = 0; i < m; ++i) any_rng = any_rng || has_rng[i];
        if (any_rng) {
            buf += "RANGES\n";
            for (int32_t i = 0; i < m; ++i) {
                if (!has_rng[i]) continue;
                buf += "    RNG  "; buf += rn[i]; buf += "  "; buf += g17(rng[i]); buf += '\n';
                flush(false);
            }
        }
        buf += "BOUNDS\n";
        for (int32_t j = 0; j < n; ++j) {
            const double lo = lp->l[j], hi = lp->u[j];
            auto line = [&](const char* t, const double* v) {
                buf += ' '; buf += t; buf += " BND  "; buf += cn[j];
                if (v) { buf += "  "; buf += g17(*v); }
                buf += '\n';
            };
            if (lo == 0.0 && !std::signbit(lo) && hi == kInf) continue;
            if (lo == hi) { line("FX", &lo); continue; }
            if (lo == -kInf && hi == kInf) { line("FR", nullptr); continue; }
            if (lo == -kInf) line("MI", nullptr);
            else if (!(lo == 0.0 && !std::signbit(lo)) || hi < 0) line("LO", &lo);
            if (hi != kInf) line("UP", &hi);
            flush(false);
        }
        buf += "ENDATA\n";
        flush(true);
        if (std::fclose(f) != 0) throw std::runtime_error(std::string("write failed: ") + path);
        return SUPR_OK;
    } catch (const std::invalid_argument& e) {
        supr::set_error(e.what());
        return SUPR_EINVAL;
    } catch (const std::bad_alloc&) {
        supr::set_error("out of host memory");
        return SUPR_ENOMEM;
    } catch (const std::exception& e) {
        supr::set_error(e.what());
        return SUPR_EINTERNAL;
    }
}

}  // extern "C"
