// Search patterns (host side): validation, probabilities, and the regular-expression
// compiler that turns a pattern into the DFA over the 64 base64 values that the GPU
// matcher walks (the reference's Pattern and Dfa, src/model/pattern.rs:57-262 and
// src/model/pattern/dfa.rs:14-200, whose regex side is built on regex-automata).
//
// Regex semantics follow the reference's configuration of regex-syntax: byte-oriented
// (no Unicode), ASCII case folding when case-insensitive, `^` / `\A` at the start of the
// 43-character hay b64[25..68], `$` / `\z` at its end, `\b` / `\B` with ASCII word
// characters. The compiler is written from those rules:
//   parse -> AST -> Thompson NFA (char sets as 64-bit masks over the base64 values) ->
//   subset construction over the 64 values, with the look-around context (start / word /
//   non-word character before, end / word / non-word after) resolved at each step.
// A DFA transition goes to ACCEPT (0, absorbing) when the epsilon closure taken before
// consuming the next character already contains the match state; column 64 ("eoi") says
// whether a match is complete at the end of the hay. States are numbered in breadth-first
// order from START = 1, as the reference does (dfa.rs:76-103). The numbering of states
// beyond START is an artifact of the construction and need not equal regex-automata's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/meikey_hip.h"
#include "mk_common.h"

namespace {

const char kAlphabet[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
constexpr u32 kMaxStates = 1u << 14;  // dfa.rs:20 MAX_STATES
constexpr size_t kMaxNfa = 1u << 15;
// Every pattern runs over the 43-character hay, so at most 43 copies of a repeated part
// consume a character; any others match empty at a position where any number of copies
// can. x{n,m} therefore matches the same hays as x{n,m} with both counts cut to 44, and
// with m >= 44 as x{min(n, 44),}: a count of 100000 builds 44 copies, not 100000.
constexpr int kRepeatCap = 44;
constexpr int kHayLen = 43;           // b64[25..68]
constexpr int kFreeLen = 42;          // Pattern::MAX_LEN (pattern.rs:21, 112)

int value_of(unsigned char c) {
    const char* p = c ? strchr(kAlphabet, c) : nullptr;
    return p ? (int)(p - kAlphabet) : -1;
}

bool is_word_value(int v) { return v < 62; }  // A-Z a-z 0-9 are word characters; + / are not

struct regex_error {
    std::string msg;
};

// ------------------------------------------------------------------ AST

enum NodeKind { N_SET, N_EMPTY, N_CONCAT, N_ALT, N_REPEAT, N_ASSERT };
enum AssertKind { A_START, A_END, A_WORD, A_NOT_WORD };

struct Node {
    NodeKind kind = N_EMPTY;
    u64 set = 0;  // N_SET: the base64 values matched (bytes outside the alphabet never occur)
    std::vector<int> kids;
    int min = 0, max = 0;  // N_REPEAT, max < 0 = unbounded
    AssertKind assert_kind = A_START;
};

struct Flags {
    bool icase = false, verbose = false;
};

// A byte set over all 256 bytes, reduced to base64 values at the end.
struct ByteSet {
    bool b[256] = {};
    void add(int lo, int hi) {
        for (int c = lo; c <= hi; c++) b[c] = true;
    }
    void negate() {
        for (bool& x : b) x = !x;
    }
    void fold() {
        for (int c = 'a'; c <= 'z'; c++)
            if (b[c] || b[c - 32]) b[c] = b[c - 32] = true;
    }
    u64 values() const {
        u64 m = 0;
        for (int v = 0; v < 64; v++)
            if (b[(unsigned char)kAlphabet[v]]) m |= 1ull << v;
        return m;
    }
};

class Parser {
   public:
    Parser(const std::string& s, bool icase) : s_(s) { flags_.icase = icase; }

    int parse(std::vector<Node>& nodes) {
        nodes_ = &nodes;
        int root = alternation();
        if (pos_ < s_.size()) {
            if (s_[pos_] == ')') fail("unopened group");
            fail("unexpected character");
        }
        return root;
    }

   private:
    const std::string& s_;
    size_t pos_ = 0;
    Flags flags_;
    std::vector<Node>* nodes_ = nullptr;
    int depth_ = 0;

    [[noreturn]] void fail(const char* what) {
        char buf[256];
        snprintf(buf, sizeof buf, "regex parse error at offset %zu: %s", pos_, what);
        throw regex_error{buf};
    }

    int add(Node n) {
        nodes_->push_back(std::move(n));
        return (int)nodes_->size() - 1;
    }

    bool eof() const { return pos_ >= s_.size(); }
    char peek() const { return s_[pos_]; }

    void skip_verbose() {
        if (!flags_.verbose) return;
        while (!eof()) {
            char c = peek();
            if (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v') {
                pos_++;
            } else if (c == '#') {
                while (!eof() && peek() != '\n') pos_++;
            } else {
                break;
            }
        }
    }

    int alternation() {
        std::vector<int> alts{concat()};
        while (!eof() && peek() == '|') {
            pos_++;
            alts.push_back(concat());
        }
        if (alts.size() == 1) return alts[0];
        Node n;
        n.kind = N_ALT;
        n.kids = alts;
        return add(n);
    }

    int concat() {
        std::vector<int> items;
        for (;;) {
            skip_verbose();
            if (eof() || peek() == '|' || peek() == ')') break;
            int a = atom();
            if (a == -2) continue;  // a flag group that only changed flags
            items.push_back(repeat(a));
        }
        if (items.empty()) return add(Node());
        if (items.size() == 1) return items[0];
        Node n;
        n.kind = N_CONCAT;
        n.kids = items;
        return add(n);
    }

    bool number(int& out) {
        size_t start = pos_;
        long v = 0;
        while (!eof() && isdigit((unsigned char)peek())) {
            v = v * 10 + (peek() - '0');
            if (v > 100000) fail("repetition count too large");
            pos_++;
        }
        out = (int)v;
        return pos_ > start;
    }

    int repeat(int atom_id) {
        skip_verbose();
        if (eof()) return atom_id;
        char c = peek();
        int mn, mx;
        if (c == '*') {
            mn = 0, mx = -1;
            pos_++;
        } else if (c == '+') {
            mn = 1, mx = -1;
            pos_++;
        } else if (c == '?') {
            mn = 0, mx = 1;
            pos_++;
        } else if (c == '{') {
            size_t save = pos_;
            pos_++;
            if (!number(mn)) {
                pos_ = save;
                fail("repetition quantifier expects a valid decimal");
            }
            if (!eof() && peek() == ',') {
                pos_++;
                if (!number(mx)) mx = -1;
            } else {
                mx = mn;
            }
            if (eof() || peek() != '}') fail("unclosed counted repetition");
            pos_++;
            if (mx >= 0 && mx < mn) fail("invalid repetition range");
        } else {
            return atom_id;
        }
        if (!eof() && peek() == '?') pos_++;  // lazy: same set of matches
        skip_verbose();
        if (!eof() && (peek() == '*' || peek() == '+' || peek() == '?' || peek() == '{'))
            fail("repetition operator missing expression");
        if ((*nodes_)[atom_id].kind == N_ASSERT) {
            // a repeated assertion is the assertion or nothing
        }
        Node n;
        n.kind = N_REPEAT;
        n.kids = {atom_id};
        n.min = mn;
        n.max = mx;
        return add(n);
    }

    int set_node(ByteSet bs) {
        if (flags_.icase) bs.fold();
        Node n;
        n.kind = N_SET;
        n.set = bs.values();
        return add(n);
    }

    int assert_node(AssertKind k) {
        Node n;
        n.kind = N_ASSERT;
        n.assert_kind = k;
        return add(n);
    }

    // Perl classes \d \w \s and their negations (ASCII).
    bool perl_class(char c, ByteSet& bs) {
        ByteSet t;
        switch (c) {
            case 'd': case 'D': t.add('0', '9'); break;
            case 'w': case 'W': t.add('0', '9'); t.add('A', 'Z'); t.add('a', 'z'); t.add('_', '_'); break;
            case 's': case 'S': t.add('\t', '\r'); t.add(' ', ' '); break;
            default: return false;
        }
        if (c == 'D' || c == 'W' || c == 'S') t.negate();
        bs = t;
        return true;
    }

    int hex_escape() {
        int v = 0, n = 0;
        if (!eof() && peek() == '{') {
            pos_++;
            while (!eof() && peek() != '}') {
                if (!isxdigit((unsigned char)peek())) fail("invalid hexadecimal digit");
                v = v * 16 + (isdigit((unsigned char)peek()) ? peek() - '0' : (tolower(peek()) - 'a' + 10));
                if (v > 0xff) fail("hexadecimal escape above 0xFF");
                pos_++;
                n++;
            }
            if (eof() || n == 0) fail("invalid hexadecimal escape");
            pos_++;
            return v;
        }
        for (int i = 0; i < 2; i++) {
            if (eof() || !isxdigit((unsigned char)peek())) fail("invalid hexadecimal escape");
            v = v * 16 + (isdigit((unsigned char)peek()) ? peek() - '0' : (tolower(peek()) - 'a' + 10));
            pos_++;
        }
        return v;
    }

    // A single-byte escape (after the backslash); returns the byte or -1 when `c` is not one.
    int byte_escape(char c) {
        switch (c) {
            case 't': return '\t';
            case 'n': return '\n';
            case 'r': return '\r';
            case 'f': return '\f';
            case 'v': return '\v';
            case 'a': return 7;
            case 'x': return hex_escape();
            default: break;
        }
        if ((unsigned char)c < 128 && ispunct((unsigned char)c)) return (unsigned char)c;
        if (c == ' ') return ' ';
        return -1;
    }

    int atom() {
        char c = peek();
        if (c == '(') return group();
        if (c == '[') {
            pos_++;
            ByteSet bs = bracket();
            return set_node(bs);
        }
        pos_++;
        switch (c) {
            case '.': {
                ByteSet bs;
                bs.add(0, 255);
                bs.b['\n'] = false;
                return set_node(bs);
            }
            case '^': return assert_node(A_START);
            case '$': return assert_node(A_END);
            case '*': case '+': case '?': case '{': fail("repetition operator missing expression");
            case '\\': {
                if (eof()) fail("incomplete escape sequence");
                char e = s_[pos_++];
                ByteSet bs;
                if (perl_class(e, bs)) return set_node(bs);
                if (e == 'b') return assert_node(A_WORD);
                if (e == 'B') return assert_node(A_NOT_WORD);
                if (e == 'A') return assert_node(A_START);
                if (e == 'z') return assert_node(A_END);
                int v = byte_escape(e);
                if (v < 0) fail("unrecognized escape sequence");
                bs.add(v, v);
                return set_node(bs);
            }
            default: {
                ByteSet bs;
                bs.add((unsigned char)c, (unsigned char)c);
                return set_node(bs);
            }
        }
    }

    void flag_list(Flags& f, bool& any_colon) {
        bool neg = false, seen = false;
        for (;;) {
            if (eof()) fail("unclosed group");
            char c = peek();
            pos_++;
            if (c == ')') {
                any_colon = false;
                return;
            }
            if (c == ':') {
                any_colon = true;
                return;
            }
            if (c == '-') {
                if (neg) fail("repeated negation in flags");
                neg = true;
                continue;
            }
            seen = true;
            switch (c) {
                case 'i': f.icase = !neg; break;
                case 'x': f.verbose = !neg; break;
                case 'm': case 's': case 'U': case 'R': break;  // no newline in the hay; greed is irrelevant
                case 'u':
                    if (!neg) fail("Unicode mode is not available (patterns are ASCII)");
                    break;
                default: fail("unrecognized flag");
            }
        }
        (void)seen;
    }

    int group() {
        pos_++;  // '('
        if (++depth_ > 250) fail("nest limit exceeded");
        Flags saved = flags_;
        if (!eof() && peek() == '?') {
            pos_++;
            if (eof()) fail("unclosed group");
            char c = peek();
            if (c == 'P' || c == '<') {
                if (c == 'P') {
                    pos_++;
                    if (eof() || peek() != '<') fail("invalid capture group name");
                }
                pos_++;
                size_t st = pos_;
                while (!eof() && (isalnum((unsigned char)peek()) || peek() == '_')) pos_++;
                if (eof() || peek() != '>' || pos_ == st) fail("invalid capture group name");
                pos_++;
            } else {
                bool colon = false;
                Flags f = flags_;
                flag_list(f, colon);
                if (!colon) {
                    // (?flags) applies to the rest of the enclosing group
                    flags_ = f;
                    depth_--;
                    return -2;
                }
                flags_ = f;
            }
        }
        int inner = alternation();
        if (eof() || peek() != ')') fail("unclosed group");
        pos_++;
        flags_ = saved;
        depth_--;
        return inner;
    }

    // Bracket class after '['; handles negation, ranges, escapes, POSIX ASCII classes,
    // nested classes and the set operators && -- ~~.
    ByteSet bracket() {
        bool neg = false;
        if (!eof() && peek() == '^') {
            neg = true;
            pos_++;
        }
        ByteSet acc = class_union(true);
        while (!eof() && peek() != ']') {
            if (s_.compare(pos_, 2, "&&") == 0 || s_.compare(pos_, 2, "--") == 0 || s_.compare(pos_, 2, "~~") == 0) {
                char op = peek();
                pos_ += 2;
                ByteSet rhs = class_union(false);
                for (int i = 0; i < 256; i++) {
                    if (op == '&') acc.b[i] = acc.b[i] && rhs.b[i];
                    else if (op == '-') acc.b[i] = acc.b[i] && !rhs.b[i];
                    else acc.b[i] = acc.b[i] != rhs.b[i];
                }
            } else {
                fail("unclosed character class");
            }
        }
        if (eof()) fail("unclosed character class");
        pos_++;  // ']'
        if (flags_.icase) acc.fold();
        if (neg) acc.negate();
        return acc;
    }

    bool posix(ByteSet& bs) {
        static const char* names[] = {"alnum", "alpha", "ascii", "blank", "cntrl", "digit", "graph",
                                      "lower", "print", "punct", "space", "upper", "word", "xdigit"};
        if (s_.compare(pos_, 2, "[:") != 0) return false;
        size_t end = s_.find(":]", pos_ + 2);
        if (end == std::string::npos) return false;
        std::string name = s_.substr(pos_ + 2, end - pos_ - 2);
        bool neg = false;
        if (!name.empty() && name[0] == '^') {
            neg = true;
            name = name.substr(1);
        }
        int idx = -1;
        for (int i = 0; i < 14; i++)
            if (name == names[i]) idx = i;
        if (idx < 0) fail("invalid POSIX character class");
        ByteSet t;
        for (int c = 0; c < 128; c++) {
            bool in = false;
            switch (idx) {
                case 0: in = isalnum(c); break;
                case 1: in = isalpha(c); break;
                case 2: in = true; break;
                case 3: in = c == ' ' || c == '\t'; break;
                case 4: in = iscntrl(c); break;
                case 5: in = isdigit(c); break;
                case 6: in = isgraph(c); break;
                case 7: in = islower(c); break;
                case 8: in = isprint(c); break;
                case 9: in = ispunct(c); break;
                case 10: in = isspace(c); break;
                case 11: in = isupper(c); break;
                case 12: in = isalnum(c) || c == '_'; break;
                default: in = isxdigit(c); break;
            }
            t.b[c] = in;
        }
        if (neg) t.negate();
        bs = t;
        pos_ = end + 2;
        return true;
    }

    // One class item: a byte (or escape) possibly followed by "-byte".
    int class_byte() {
        char c = s_[pos_++];
        if (c == '\\') {
            if (eof()) fail("incomplete escape sequence");
            char e = s_[pos_++];
            int v = byte_escape(e);
            if (v < 0) fail("unrecognized escape sequence in class");
            return v;
        }
        return (unsigned char)c;
    }

    // Items until ']' or a set operator. `first` allows a leading ']' literal.
    ByteSet class_union(bool first) {
        ByteSet bs;
        bool any = false;
        while (!eof()) {
            char c = peek();
            if (c == ']' && !(first && !any)) break;
            if (any && (s_.compare(pos_, 2, "&&") == 0 || s_.compare(pos_, 2, "--") == 0 ||
                        s_.compare(pos_, 2, "~~") == 0))
                break;
            if (c == '[') {
                ByteSet t;
                if (posix(t)) {
                    for (int i = 0; i < 256; i++) bs.b[i] = bs.b[i] || t.b[i];
                } else {
                    pos_++;
                    t = bracket();
                    for (int i = 0; i < 256; i++) bs.b[i] = bs.b[i] || t.b[i];
                }
                any = true;
                continue;
            }
            if (c == '\\' && pos_ + 1 < s_.size()) {
                ByteSet t;
                if (perl_class(s_[pos_ + 1], t)) {
                    pos_ += 2;
                    for (int i = 0; i < 256; i++) bs.b[i] = bs.b[i] || t.b[i];
                    any = true;
                    continue;
                }
            }
            int lo = class_byte();
            if (!eof() && peek() == '-' && pos_ + 1 < s_.size() && s_[pos_ + 1] != ']' &&
                s_.compare(pos_, 2, "--") != 0) {
                pos_++;
                int hi = class_byte();
                if (hi < lo) fail("invalid character class range");
                bs.add(lo, hi);
            } else {
                bs.add(lo, lo);
            }
            any = true;
        }
        if (!any) fail("empty character class");
        return bs;
    }
};

// Shortest match length (Hir::minimum_len): look-arounds count zero. Saturates at
// kMinLenCap (every length above the 43-character hay is alike here), so nested counts
// never overflow: regex-syntax computes the same with checked arithmetic.
constexpr long kMinLenCap = kHayLen + 1;
long min_len(const std::vector<Node>& nodes, int id) {
    const Node& n = nodes[id];
    switch (n.kind) {
        case N_SET: return 1;
        case N_EMPTY: case N_ASSERT: return 0;
        case N_CONCAT: {
            long t = 0;
            for (int k : n.kids) t = std::min(kMinLenCap, t + min_len(nodes, k));
            return t;
        }
        case N_ALT: {
            long m = -1;
            for (int k : n.kids) {
                long v = min_len(nodes, k);
                if (m < 0 || v < m) m = v;
            }
            return m < 0 ? 0 : m;
        }
        default: return std::min(kMinLenCap, (long)n.min * min_len(nodes, n.kids[0]));
    }
}

bool has_word_assert(const std::vector<Node>& nodes) {
    for (const Node& n : nodes)
        if (n.kind == N_ASSERT && (n.assert_kind == A_WORD || n.assert_kind == A_NOT_WORD)) return true;
    return false;
}

// ------------------------------------------------------------------ NFA

enum StateKind { S_SET, S_SPLIT, S_EPS, S_ASSERT, S_MATCH };

struct NState {
    StateKind kind;
    u64 set = 0;
    int out = -1, out2 = -1;
    AssertKind assert_kind = A_START;
};

struct Frag {
    int start;
    std::vector<int*> dangling;  // pointers into NState::out fields (stable: nfa reserved)
};

class NfaBuilder {
   public:
    std::vector<NState> st;
    explicit NfaBuilder(const std::vector<Node>& nodes) : nodes_(nodes) { st.reserve(kMaxNfa + 16); }

    int build(int root) {
        Frag f = frag(root);
        int m = add({S_MATCH});
        patch(f, m);
        return f.start;
    }

   private:
    const std::vector<Node>& nodes_;

    int add(NState s) {
        if (st.size() >= kMaxNfa) throw regex_error{"too complex nfa"};
        st.push_back(s);
        return (int)st.size() - 1;
    }

    void patch(Frag& f, int to) {
        for (int* p : f.dangling) *p = to;
        f.dangling.clear();
    }

    Frag frag(int id) {
        const Node& n = nodes_[id];
        switch (n.kind) {
            case N_SET: {
                int s = add({S_SET, n.set});
                return {s, {&st[s].out}};
            }
            case N_EMPTY: {
                int s = add({S_EPS});
                return {s, {&st[s].out}};
            }
            case N_ASSERT: {
                NState a{S_ASSERT};
                a.assert_kind = n.assert_kind;
                int s = add(a);
                return {s, {&st[s].out}};
            }
            case N_CONCAT: {
                Frag f = frag(n.kids[0]);
                for (size_t i = 1; i < n.kids.size(); i++) {
                    Frag g = frag(n.kids[i]);
                    patch(f, g.start);
                    f.dangling = g.dangling;
                }
                return f;
            }
            case N_ALT: {
                // chain of splits
                Frag acc = frag(n.kids.back());
                for (size_t i = n.kids.size() - 1; i-- > 0;) {
                    Frag g = frag(n.kids[i]);
                    int s = add({S_SPLIT});
                    st[s].out = g.start;
                    st[s].out2 = acc.start;
                    Frag r{s, g.dangling};
                    r.dangling.insert(r.dangling.end(), acc.dangling.begin(), acc.dangling.end());
                    acc = r;
                }
                return acc;
            }
            default: {
                // N_REPEAT: min copies, then (max - min) optional copies or a star
                const int kid = n.kids[0];
                int rmin = n.min, rmax = n.max;
                if (rmax < 0 || rmax >= kRepeatCap) {
                    rmax = -1;
                    rmin = std::min(rmin, kRepeatCap);
                }
                int e = add({S_EPS});
                Frag f{e, {&st[e].out}};
                for (int i = 0; i < rmin; i++) {
                    Frag g = frag(kid);
                    patch(f, g.start);
                    f.dangling = g.dangling;
                }
                if (rmax < 0) {
                    Frag g = frag(kid);
                    int s = add({S_SPLIT});
                    st[s].out = g.start;
                    patch(g, s);
                    patch(f, s);
                    return {e, {&st[s].out2}};
                }
                std::vector<int*> outs;
                for (int i = rmin; i < rmax; i++) {
                    Frag g = frag(kid);
                    int s = add({S_SPLIT});
                    st[s].out = g.start;
                    patch(f, s);
                    outs.push_back(&st[s].out2);
                    f.dangling = g.dangling;
                }
                f.dangling.insert(f.dangling.end(), outs.begin(), outs.end());
                f.start = e;
                return f;
            }
        }
    }
};

// ------------------------------------------------------------------ DFA

// Context of a position: the character before (0 start, 1 word, 2 non-word) and after
// (0 end, 1 word, 2 non-word).
bool assert_holds(AssertKind k, int prev, int next) {
    const bool wp = prev == 1, wn = next == 1;
    switch (k) {
        case A_START: return prev == 0;
        case A_END: return next == 0;
        case A_WORD: return wp != wn;
        default: return wp == wn;
    }
}

struct Dfa {
    std::vector<u16> trans;  // states * 64
    std::vector<u8> eoi;
};

// Subset construction. DFA state = (NFA set-states waiting to consume a character,
// context of the character before). `words` says whether \b or \B occur; if not, the
// word / non-word context is merged.
//
// The closure of a state depends on the value read only through that value's context
// (word or non-word, one class when `words` is false), so it is computed once per
// context and the 64 transitions are read off it. The work is bounded for untrusted
// pattern text like regex-automata's determinize_size_limit (the reference's 16 MiB
// BUILD_LIMIT, src/model/pattern/dfa.rs:23, 64-65): the bytes of the state keys kept
// and the closure steps taken are counted, and past the limit the pattern is too complex.
constexpr size_t kBuildLimit = 16u << 20;        // bytes of DFA state keys
constexpr u64 kClosureSteps = 1ull << 25;        // NFA states visited, summed over closures
constexpr size_t kStateKeyOverhead = 64;         // map node and vector header per state

Dfa determinize(const std::vector<NState>& nfa, int start, bool words) {
    std::vector<int> stack;
    std::vector<u32> mark(nfa.size(), 0);
    u32 stamp = 0;
    u64 steps = 0;
    size_t key_bytes = 0;
    // closure of `seeds` plus the start state in context (prev, next): SET states and
    // whether MATCH is reached
    auto closure = [&](const std::vector<int>& seeds, int prev, int next, std::vector<int>& sets) {
        stamp++;
        sets.clear();
        bool match = false;
        stack.assign(seeds.begin(), seeds.end());
        stack.push_back(start);
        while (!stack.empty()) {
            int s = stack.back();
            stack.pop_back();
            if (s < 0 || mark[s] == stamp) continue;
            mark[s] = stamp;
            if (++steps > kClosureSteps) throw regex_error{"too complex build"};
            const NState& n = nfa[s];
            switch (n.kind) {
                case S_SET: sets.push_back(s); break;
                case S_MATCH: match = true; break;
                case S_EPS: stack.push_back(n.out); break;
                case S_SPLIT:
                    stack.push_back(n.out2);
                    stack.push_back(n.out);
                    break;
                case S_ASSERT:
                    if (assert_holds(n.assert_kind, prev, next)) stack.push_back(n.out);
                    break;
            }
        }
        return match;
    };
    auto ctx_of = [&](int v) { return words ? (is_word_value(v) ? 1 : 2) : 1; };
    // the values of each next-character context (1 word, 2 non-word)
    u64 ctx_values[3] = {0, 0, 0};
    for (int v = 0; v < 64; v++) ctx_values[ctx_of(v)] |= 1ull << v;

    Dfa d;
    std::map<std::vector<int>, u32> ids;  // key: sorted waiting states + [prev context]
    std::vector<std::vector<int>> order;
    // state 0 = ACCEPT: every value stays in ACCEPT, end of input matches
    d.trans.assign(64, 0);
    d.eoi.push_back(1);
    std::vector<int> first = {0};  // no waiting states, prev = start
    ids[first] = 1;
    order.push_back(first);
    std::vector<int> sets, key;
    std::vector<int> outs[64];
    for (size_t i = 0; i < order.size(); i++) {
        const std::vector<int> cur = order[i];
        const int prev = cur.back();
        const std::vector<int> waiting(cur.begin(), cur.end() - 1);
        u16 row[64];
        for (int ctx = 1; ctx <= 2; ctx++) {
            const u64 vals = ctx_values[ctx];
            if (vals == 0) continue;
            if (closure(waiting, prev, ctx, sets)) {
                for (int v = 0; v < 64; v++)
                    if ((vals >> v) & 1) row[v] = 0;
                continue;
            }
            for (int v = 0; v < 64; v++) outs[v].clear();
            for (int s : sets) {
                u64 m = nfa[s].set & vals;
                steps += (u64)__builtin_popcountll(m);
                while (m) {
                    const int v = __builtin_ctzll(m);
                    m &= m - 1;
                    outs[v].push_back(nfa[s].out);
                }
            }
            if (steps > kClosureSteps) throw regex_error{"too complex build"};
            for (int v = 0; v < 64; v++) {
                if (!((vals >> v) & 1)) continue;
                key = outs[v];
                std::sort(key.begin(), key.end());
                key.erase(std::unique(key.begin(), key.end()), key.end());
                key.push_back(ctx);
                auto it = ids.find(key);
                if (it == ids.end()) {
                    if (order.size() + 1 >= kMaxStates) throw regex_error{"too complex"};
                    key_bytes += key.size() * sizeof(int) + kStateKeyOverhead;
                    if (key_bytes > kBuildLimit) throw regex_error{"too complex build"};
                    it = ids.emplace(key, (u32)order.size() + 1).first;
                    order.push_back(key);
                }
                row[v] = (u16)it->second;
            }
        }
        d.trans.insert(d.trans.end(), row, row + 64);
        d.eoi.push_back(closure(waiting, prev, 0, sets) ? 1 : 0);
    }
    return d;
}

// dfa.rs:164-185: the first hay character is one of A..P (values 0..15, 1/16 each), the
// others uniform over the 64 values; the probability of ending in ACCEPT or a state that
// accepts at the end.
double dfa_probability(const Dfa& d) {
    const size_t n = d.eoi.size();
    std::vector<double> p(n, 0.0), q(n, 0.0);
    p[1] = 1.0;
    for (int i = 0; i < kHayLen; i++) {
        const int choices = i == 0 ? 16 : 64;
        const double w = 1.0 / choices;
        std::fill(q.begin(), q.end(), 0.0);
        for (size_t s = 0; s < n; s++) {
            if (p[s] == 0.0) continue;
            for (int v = 0; v < choices; v++) q[d.trans[s * 64 + v]] += p[s] * w;
        }
        std::swap(p, q);
    }
    double ends = 0;
    for (size_t s = 1; s < n; s++)
        if (d.eoi[s]) ends += p[s];
    return std::min(1.0, p[0] + ends);
}

thread_local std::string g_pattern_error;

mk_status perr(int* code, int c, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_pattern_error = buf;
    if (code) *code = c;
    return MK_ERR_INVALID;
}

}  // namespace

struct mk_pattern {
    std::string text;
    int case_sensitive = 1;
    u32 mode = MK_MODE_CONTAINS;
    std::vector<u8> needle;  // Needle bytes: lower-cased text, or the DFA table
    Dfa dfa;
    double probability = 0;
};

extern "C" {

void mk_pattern_last_error(char* buf, size_t len) {
    if (buf && len) snprintf(buf, len, "%s", g_pattern_error.c_str());
}

mk_status mk_pattern_new(const char* text, uint32_t text_len, int case_sensitive, uint32_t mode, mk_pattern** out,
                         int* error_kind) {
    if (error_kind) *error_kind = MK_PATTERN_OK;
    if (!out || (text_len && !text)) return perr(error_kind, MK_PATTERN_OK, "null pointer");
    *out = nullptr;
    const std::string s(text ? text : "", text_len);
    if (s.empty()) return perr(error_kind, MK_PATTERN_EMPTY, "the text is empty");
    mk_pattern* p = new mk_pattern();
    p->text = s;
    p->case_sensitive = case_sensitive ? 1 : 0;
    p->mode = mode;
    if (mode == MK_MODE_REGEX) {
        std::vector<Node> nodes;
        try {
            Parser parser(s, !case_sensitive);
            int root = parser.parse(nodes);
            if (min_len(nodes, root) == 0) {
                delete p;
                return perr(error_kind, MK_PATTERN_MATCHES_EMPTY,
                            "a regular expression that matches the empty string cannot be used (it matches every key)");
            }
            NfaBuilder nb(nodes);
            int start = nb.build(root);
            p->dfa = determinize(nb.st, start, has_word_assert(nodes));
        } catch (const regex_error& e) {
            delete p;
            if (e.msg == "too complex")
                return perr(error_kind, MK_PATTERN_TOO_COMPLEX,
                            "the regular expression is too complex (the DFA has more than %u states)", kMaxStates);
            if (e.msg == "too complex nfa")
                return perr(error_kind, MK_PATTERN_TOO_COMPLEX,
                            "the regular expression is too complex (its automaton has more than %zu states)", kMaxNfa);
            if (e.msg == "too complex build")
                return perr(error_kind, MK_PATTERN_TOO_COMPLEX,
                            "the regular expression is too complex (building its DFA exceeds the %zu MiB limit)",
                            kBuildLimit >> 20);
            return perr(error_kind, MK_PATTERN_REGEX, "invalid regular expression: %s", e.msg.c_str());
        }
        p->probability = dfa_probability(p->dfa);
        if (p->probability == 0.0) {
            delete p;
            return perr(error_kind, MK_PATTERN_NEVER_MATCHES,
                        "no public key matches this regular expression (the characters are A-Z a-z 0-9 + /, and the "
                        "one right after the fixed part is A-P)");
        }
        // Dfa::gpu_table (dfa.rs:187-200): per state 64 transitions and the eoi flag, u16 LE
        const size_t n = p->dfa.eoi.size();
        p->needle.resize(n * 65 * 2);
        for (size_t st = 0; st < n; st++)
            for (int c = 0; c < 65; c++) {
                const u16 v = c < 64 ? p->dfa.trans[st * 64 + c] : p->dfa.eoi[st];
                p->needle[(st * 65 + c) * 2] = (u8)v;
                p->needle[(st * 65 + c) * 2 + 1] = (u8)(v >> 8);
            }
        *out = p;
        return MK_OK;
    }
    if (mode != MK_MODE_CONTAINS && mode != MK_MODE_SUFFIX) {
        delete p;
        return perr(error_kind, MK_PATTERN_OK, "unknown mode %u", mode);
    }
    for (char c : s)
        if (value_of((unsigned char)c) < 0) {
            delete p;
            return perr(error_kind, MK_PATTERN_INVALID_CHAR, "invalid character: '%c' (allowed: A-Z a-z 0-9 + /)", c);
        }
    if (s.size() > (size_t)kFreeLen) {
        delete p;
        return perr(error_kind, MK_PATTERN_TOO_LONG, "the text is too long: %zu characters (at most %d)", s.size(),
                    kFreeLen);
    }
    // Needle::new (gpu.rs:646-649): lower-cased when case-insensitive
    for (char c : s) p->needle.push_back((u8)(case_sensitive ? c : tolower((unsigned char)c)));
    // Pattern::probability (pattern.rs:224-248, 252-258)
    auto per_char = [&](char c) { return !case_sensitive && isalpha((unsigned char)c) ? 2.0 / 64.0 : 1.0 / 64.0; };
    if (mode == MK_MODE_SUFFIX) {
        double pr = 1.0;
        for (char c : s) pr *= per_char(c);
        p->probability = pr;
    } else {
        double rest = 1.0;
        for (size_t i = 1; i < s.size(); i++) rest *= per_char(s[i]);
        const double positions = (double)(kFreeLen + 1 - (int)s.size());
        const double free = positions * per_char(s[0]) * rest;
        const char head = case_sensitive ? s[0] : (char)toupper((unsigned char)s[0]);
        const double at_start = (head >= 'A' && head <= 'P') ? rest / 16.0 : 0.0;
        p->probability = free + at_start;
    }
    *out = p;
    return MK_OK;
}

mk_status mk_pattern_free(mk_pattern* p) {
    delete p;
    return MK_OK;
}

mk_status mk_pattern_needle(const mk_pattern* p, const uint8_t** bytes, uint32_t* len, uint32_t* fold, uint32_t* mode) {
    if (!p || !bytes || !len || !fold || !mode) return MK_ERR_INVALID;
    *bytes = p->needle.data();
    *mode = p->mode;
    if (p->mode == MK_MODE_REGEX) {
        *len = (uint32_t)p->dfa.eoi.size();
        *fold = 0;
    } else {
        *len = (uint32_t)p->needle.size();
        *fold = p->case_sensitive ? 0 : 1;
    }
    return MK_OK;
}

size_t mk_pattern_needle_bytes(const mk_pattern* p) { return p ? p->needle.size() : 0; }

double mk_pattern_probability(const mk_pattern* p) { return p ? p->probability : 0.0; }

uint32_t mk_pattern_dfa_states(const mk_pattern* p) { return p && p->mode == MK_MODE_REGEX ? (uint32_t)p->dfa.eoi.size() : 0; }

int mk_pattern_is_match(const mk_pattern* p, const char* b64) {
    if (!p || !b64) return -1;
    const char* hay = b64 + 25;
    if (p->mode == MK_MODE_REGEX) {
        u32 s = MK_DFA_START;
        for (int i = 0; i < kHayLen; i++) {
            const int v = value_of((unsigned char)hay[i]);
            s = p->dfa.trans[s * 64 + (v < 0 ? 0 : v)];
            if (s == MK_DFA_ACCEPT) return 1;
        }
        return p->dfa.eoi[s] ? 1 : 0;
    }
    const size_t n = p->text.size();
    auto same = [&](const char* w) {
        for (size_t i = 0; i < n; i++) {
            const char a = w[i], b = p->text[i];
            if (p->case_sensitive ? a != b : tolower((unsigned char)a) != tolower((unsigned char)b)) return false;
        }
        return true;
    };
    if (p->mode == MK_MODE_SUFFIX) return same(b64 + 68 - n) ? 1 : 0;
    for (size_t st = 0; st + n <= (size_t)kHayLen; st++)
        if (same(hay + st)) return 1;
    return 0;
}

}  // extern "C"
