// ORACLE — test infrastructure only. Nothing in the product links, imports or runs this file.
//
// Tree-walking restatement of pygranite's expression language, used by the CPU oracle to
// evaluate `AdditionalCompute` programs.
//
//   grammar ............ source/granite/TrajectoryIntegrator.cpp:10-26 (cpp-peglib PEG)
//   semantic actions ... source/granite/TrajectoryIntegrator.cpp:35-135
//   identifier binding . source/granite/TrajectoryIntegrator.cpp:744-768 (replace_environment)
//   validation ......... source/granite/TrajectoryIntegrator.tpp:653-697
//   evaluation ......... source/granite/TrajectoryIntegrator.cpp:695-742 (host interpret),
//                        source/granite/TrajectoryIntegrator.hpp:539-627 (device interpret)
//   node type numbers .. source/granite/Language.hpp:17-73
//
// The parser is a hand-written recursive descent that mirrors the PEG's ordered choice:
// '(' Expr ')' / Function / Register / OldRegister / Identifier / Constant, with `+ -` and
// `* /` as two left-associative precedence levels. A Register literal commits (PEG does not
// backtrack into Identifier once Register matched), so `velocity` fails to parse exactly
// like it does under cpp-peglib. `abs(...)` is accepted only when `allow_abs` is set (an
// extension that the reference grammar does not have).
#pragma once

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle
{

enum NodeType : int
{
    N_X = 1, N_Y = 2, N_Z = 3, N_CV = 4,
    N_F0 = 5,  // .. N_F7 = 12
    N_C0 = 13, // .. N_C3 = 16
    N_CONST = 17,
    N_ADD = 18, N_MUL = 19, N_SUB = 20, N_DIV = 21,
    N_SQRT = 22, N_EXP = 23, N_LN = 24, N_SIN = 25, N_COS = 26, N_TAN = 27,
    N_ASIN = 28, N_ACOS = 29, N_ATAN = 30, N_ATAN2 = 31,
    N_U = 33, N_V = 34, N_W = 35,
    N_ABS = 40 // extension, not in the reference grammar
};

struct ENode
{
    int type{0};
    float value{0.f};
    int left{-1}, right{-1};
};

struct Program
{
    std::vector<ENode> nodes; // nodes[root] is the root
    int root{-1};
};

class Parser
{
public:
    Parser(const std::string & src, const std::map<std::string, int> & env, bool allow_abs)
        : s_(src), env_(env), allow_abs_(allow_abs)
    {}

    Program parse()
    {
        Program p;
        prog_ = &p;
        int r = expr(1);
        ws();
        if (pos_ != s_.size()) fail("syntax error, unexpected '" + s_.substr(pos_, 1) + "'");
        p.root = r;
        return p;
    }

private:
    const std::string & s_;
    const std::map<std::string, int> & env_;
    bool allow_abs_;
    size_t pos_{0};
    Program * prog_{nullptr};

    [[noreturn]] void fail(const std::string & msg) { throw std::runtime_error(msg); }

    void ws()
    {
        while (pos_ < s_.size() && (s_[pos_] == ' ' || s_[pos_] == '\t')) ++pos_;
    }

    int add(int type, float value, int l, int r)
    {
        prog_->nodes.push_back(ENode{type, value, l, r});
        return static_cast<int>(prog_->nodes.size()) - 1;
    }

    static int prec(char c) { return (c == '+' || c == '-') ? 1 : 2; }

    int expr(int min_prec)
    {
        int lhs = primary();
        for (;;)
        {
            ws();
            if (pos_ >= s_.size()) break;
            char c = s_[pos_];
            if (c != '+' && c != '-' && c != '*' && c != '/') break;
            int p = prec(c);
            if (p < min_prec) break;
            ++pos_;
            int rhs = expr(p + 1);
            int t = c == '+' ? N_ADD : c == '-' ? N_SUB : c == '*' ? N_MUL : N_DIV;
            lhs = add(t, 0.f, lhs, rhs);
        }
        return lhs;
    }

    bool lit(const char * w)
    {
        size_t n = std::char_traits<char>::length(w);
        if (s_.compare(pos_, n, w) == 0)
        {
            pos_ += n;
            return true;
        }
        return false;
    }

    // Function alternatives in the grammar's order; `name` then '(' (whitespace allowed).
    bool function(int & out)
    {
        struct F { const char * name; int type; bool two; };
        static const F fns[] = {{"sqrt", N_SQRT, false}, {"exp", N_EXP, false},
                                {"ln", N_LN, false},     {"sin", N_SIN, false},
                                {"cos", N_COS, false},   {"tan", N_TAN, false},
                                {"asin", N_ASIN, false}, {"acos", N_ACOS, false},
                                {"atan", N_ATAN, false}, {"atan2", N_ATAN2, true},
                                {"abs", N_ABS, false}};
        const size_t start = pos_;
        for (const F & f : fns)
        {
            if (f.type == N_ABS && !allow_abs_) continue;
            pos_ = start;
            if (!lit(f.name)) continue;
            ws();
            if (pos_ >= s_.size() || s_[pos_] != '(') continue;
            ++pos_;
            int a = expr(1);
            int b = -1;
            ws();
            if (f.two)
            {
                if (pos_ >= s_.size() || s_[pos_] != ',') fail("syntax error, expecting ','");
                ++pos_;
                b = expr(1);
                ws();
            }
            if (pos_ >= s_.size() || s_[pos_] != ')') fail("syntax error, expecting ')'");
            ++pos_;
            out = add(f.type, 0.f, a, b);
            return true;
        }
        pos_ = start;
        return false;
    }

    int primary()
    {
        ws();
        if (pos_ >= s_.size()) fail("syntax error, unexpected end of input");
        char c = s_[pos_];
        if (c == '(')
        {
            ++pos_;
            int e = expr(1);
            ws();
            if (pos_ >= s_.size() || s_[pos_] != ')') fail("syntax error, expecting ')'");
            ++pos_;
            return e;
        }
        int f;
        if (function(f)) return f;
        // Register <- 'x' | 'y' | 'z' | 'cv' | 'u' | 'v' | 'w'
        if (s_.compare(pos_, 2, "cv") == 0)
        {
            pos_ += 2;
            return add(N_CV, 0.f, -1, -1);
        }
        switch (c)
        {
            case 'x': ++pos_; return add(N_X, 0.f, -1, -1);
            case 'y': ++pos_; return add(N_Y, 0.f, -1, -1);
            case 'z': ++pos_; return add(N_Z, 0.f, -1, -1);
            case 'u': ++pos_; return add(N_U, 0.f, -1, -1);
            case 'v': ++pos_; return add(N_V, 0.f, -1, -1);
            case 'w': ++pos_; return add(N_W, 0.f, -1, -1);
            default: break;
        }
        // OldRegister <- 'f0'..'f7' | 'c0'..'c3'
        if (pos_ + 1 < s_.size() + 0 && ((c == 'f' && s_[pos_ + 1] >= '0' && s_[pos_ + 1] <= '7') ||
                                         (c == 'c' && s_[pos_ + 1] >= '0' && s_[pos_ + 1] <= '3')))
            fail("Identifier '" + s_.substr(pos_, 2) + "' is no longer allowed");
        // Identifier <- [_a-zA-Z][a-zA-Z0-9_]*
        if (c == '_' || std::isalpha(static_cast<unsigned char>(c)))
        {
            size_t b = pos_;
            while (pos_ < s_.size() &&
                   (s_[pos_] == '_' || std::isalnum(static_cast<unsigned char>(s_[pos_]))))
                ++pos_;
            std::string id = s_.substr(b, pos_ - b);
            auto it = env_.find(id);
            if (it == env_.end()) fail("Invalid identifier '" + id + "'.");
            return add(it->second, 0.f, -1, -1);
        }
        // Constant <- [-]?[0-9]+[.][0-9]+ / [-]?[0-9]+
        {
            size_t b = pos_;
            size_t q = pos_;
            if (q < s_.size() && s_[q] == '-') ++q;
            size_t d0 = q;
            while (q < s_.size() && std::isdigit(static_cast<unsigned char>(s_[q]))) ++q;
            if (q > d0)
            {
                if (q + 1 < s_.size() && s_[q] == '.' &&
                    std::isdigit(static_cast<unsigned char>(s_[q + 1])))
                {
                    ++q;
                    while (q < s_.size() && std::isdigit(static_cast<unsigned char>(s_[q]))) ++q;
                }
                pos_ = q;
                std::string tok = s_.substr(b, q - b);
                return add(N_CONST, std::strtof(tok.c_str(), nullptr), -1, -1);
            }
        }
        fail("syntax error, unexpected '" + s_.substr(pos_, 1) + "'");
    }
};

inline bool contains(const Program & p, int type)
{
    for (const ENode & n : p.nodes)
        if (n.type == type) return true;
    return false;
}

// Evaluation over the 19-slot reference vector laid out as in
// source/granite/TrajectoryIntegrator.tpp:381-439:
// [0..2] new position, [3] cv, [4..11] volumes F0..F7, [12..15] constants C0..C3,
// [16..18] velocity (u, v, w).
template <class R> R eval(const Program & p, int i, const R * ref)
{
    if (i < 0) return R(0);
    const ENode & n = p.nodes[static_cast<size_t>(i)];
    switch (n.type)
    {
        case N_X: return ref[0];
        case N_Y: return ref[1];
        case N_Z: return ref[2];
        case N_CV: return ref[3];
        case N_U: return ref[16];
        case N_V: return ref[17];
        case N_W: return ref[18];
        case N_CONST: return R(n.value);
        case N_ADD: return eval(p, n.left, ref) + eval(p, n.right, ref);
        case N_MUL: return eval(p, n.left, ref) * eval(p, n.right, ref);
        case N_SUB: return eval(p, n.left, ref) - eval(p, n.right, ref);
        case N_DIV: return eval(p, n.left, ref) / eval(p, n.right, ref);
        case N_SQRT: return std::sqrt(eval(p, n.left, ref));
        case N_EXP: return std::exp(eval(p, n.left, ref));
        case N_LN: return std::log(eval(p, n.left, ref));
        case N_SIN: return std::sin(eval(p, n.left, ref));
        case N_COS: return std::cos(eval(p, n.left, ref));
        case N_TAN: return std::tan(eval(p, n.left, ref));
        case N_ASIN: return std::asin(eval(p, n.left, ref));
        case N_ACOS: return std::acos(eval(p, n.left, ref));
        case N_ATAN: return std::atan(eval(p, n.left, ref));
        case N_ATAN2: return std::atan2(eval(p, n.left, ref), eval(p, n.right, ref));
        case N_ABS: return std::fabs(eval(p, n.left, ref));
        default:
            if (n.type >= N_F0 && n.type <= N_F0 + 7) return ref[4 + (n.type - N_F0)];
            if (n.type >= N_C0 && n.type <= N_C0 + 3) return ref[12 + (n.type - N_C0)];
            return R(0);
    }
}

} // namespace oracle
