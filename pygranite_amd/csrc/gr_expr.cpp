// Expression compiler: AdditionalCompute source -> postfix bytecode for the step kernel's
// wavefront-uniform interpreter.
//
// Language (source/granite/TrajectoryIntegrator.cpp:10-26): binary + - (lowest) and * /, all
// left-associative; functions sqrt exp ln sin cos tan asin acos atan atan2(a, b); registers
// x y z cv u v w; loader names bound to F0..F7 (volumes) and C0..C3 (constants); constants
// -?[0-9]+(.[0-9]+)?. The old register names f0..f7, c0..c3 are rejected
// (cpp:110-114), unknown identifiers are rejected (cpp:744-768), and the 2D / curvature checks
// of TrajectoryIntegrator.tpp:661-696 apply. The choice order of the grammar's
// PrimaryExpression is kept: a register letter is taken as the register even when letters
// follow, which makes e.g. `velocity` a syntax error exactly as under the PEG parser.
//
// Code is emitted directly while parsing (operands first, then the operator), so the
// bytecode is the postfix form of the same tree the reference flattens in pre-order
// (cpp:770-798).
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "gr_internal.h"

namespace
{

struct Env
{
    const gr_expr_env * e;
};

class Compiler
{
public:
    Compiler(const char * src, const gr_expr_env & env, gr_program & out)
        : src_(src), len_(std::strlen(src)), env_(env), out_(out)
    {}

    bool run(std::string & err)
    {
        out_.n = 0;
        out_.depth = 0;
        depth_ = 0;
        nimm_ = 0;
        if (!expr(1)) return fail_out(err);
        skip_ws();
        if (pos_ != len_)
        {
            msg_ = std::string("syntax error, unexpected '") + src_[pos_] + "'";
            return fail_out(err);
        }
        return true;
    }

private:
    const char * src_;
    size_t len_;
    size_t pos_{0};
    const gr_expr_env & env_;
    gr_program & out_;
    uint32_t depth_{0};
    uint32_t nimm_{0};
    std::string msg_;

    bool fail_out(std::string & err)
    {
        err = msg_.empty() ? "syntax error" : msg_;
        return false;
    }

    bool fail(const std::string & m)
    {
        if (msg_.empty()) msg_ = m;
        return false;
    }

    void skip_ws()
    {
        while (pos_ < len_ && (src_[pos_] == ' ' || src_[pos_] == '\t')) ++pos_;
    }

    bool emit(uint32_t op, uint32_t arg, int stack_delta)
    {
        if (out_.n >= GR_MAX_CODE) return fail("expression too long");
        out_.code[out_.n++] = op | (arg << 8);
        int d = int(depth_) + stack_delta;
        depth_ = uint32_t(d);
        if (depth_ > out_.depth) out_.depth = depth_;
        if (out_.depth > GR_MAX_STACK) return fail("expression nests too deeply");
        return true;
    }

    bool push_ref(uint32_t slot) { return emit(gr::OP_REF, slot, +1); }

    bool push_const(float v)
    {
        if (nimm_ >= GR_MAX_CODE) return fail("expression too long");
        out_.imm[nimm_] = v;
        return emit(gr::OP_CONST, nimm_++, +1);
    }

    bool expr(int min_level)
    {
        if (!primary()) return false;
        for (;;)
        {
            skip_ws();
            if (pos_ >= len_) return true;
            const char c = src_[pos_];
            int level;
            uint32_t op;
            switch (c)
            {
                case '+': level = 1; op = gr::OP_ADD; break;
                case '-': level = 1; op = gr::OP_SUB; break;
                case '*': level = 2; op = gr::OP_MUL; break;
                case '/': level = 2; op = gr::OP_DIV; break;
                default: return true;
            }
            if (level < min_level) return true;
            ++pos_;
            if (!expr(level + 1)) return false;
            if (!emit(op, 0, -1)) return false;
        }
    }

    bool starts(const char * w) const
    {
        const size_t n = std::strlen(w);
        return pos_ + n <= len_ && std::strncmp(src_ + pos_, w, n) == 0;
    }

    // 'name' '(' Expression (',' Expression)? ')' in the grammar's choice order
    int function()
    {
        struct F
        {
            const char * name;
            uint32_t op;
            bool two;
        };
        static const F fns[] = {{"sqrt", gr::OP_SQRT, false}, {"exp", gr::OP_EXP, false},
                                {"ln", gr::OP_LN, false},     {"sin", gr::OP_SIN, false},
                                {"cos", gr::OP_COS, false},   {"tan", gr::OP_TAN, false},
                                {"asin", gr::OP_ASIN, false}, {"acos", gr::OP_ACOS, false},
                                {"atan", gr::OP_ATAN, false}, {"atan2", gr::OP_ATAN2, true},
                                {"abs", gr::OP_ABS, false}};
        const size_t start = pos_;
        for (const F & f : fns)
        {
            if (f.op == gr::OP_ABS && !(env_.flags & GR_EXPR_ALLOW_ABS)) continue;
            pos_ = start;
            if (!starts(f.name)) continue;
            pos_ += std::strlen(f.name);
            skip_ws();
            if (pos_ >= len_ || src_[pos_] != '(') continue;
            ++pos_;
            if (!expr(1)) return -1;
            skip_ws();
            if (f.two)
            {
                if (pos_ >= len_ || src_[pos_] != ',') return fail("syntax error, expecting ','") ? 1 : -1;
                ++pos_;
                if (!expr(1)) return -1;
                skip_ws();
            }
            if (pos_ >= len_ || src_[pos_] != ')') return fail("syntax error, expecting ')'") ? 1 : -1;
            ++pos_;
            return emit(f.op, 0, f.two ? -1 : 0) ? 1 : -1;
        }
        pos_ = start;
        return 0;
    }

    bool primary()
    {
        skip_ws();
        if (pos_ >= len_) return fail("syntax error, unexpected end of input");
        const char c = src_[pos_];
        if (c == '(')
        {
            ++pos_;
            if (!expr(1)) return false;
            skip_ws();
            if (pos_ >= len_ || src_[pos_] != ')') return fail("syntax error, expecting ')'");
            ++pos_;
            return true;
        }
        const int f = function();
        if (f < 0) return false;
        if (f > 0) return true;

        // registers (ref slots of TrajectoryIntegrator.tpp:381-439)
        if (starts("cv"))
        {
            if (env_.curvature != GR_CURV_INDIVIDUAL_AND_TOTAL)
                return fail("Compute code contains undefinded token 'cv'");
            pos_ += 2;
            return push_ref(3);
        }
        switch (c)
        {
            case 'x': ++pos_; return push_ref(0);
            case 'y': ++pos_; return push_ref(1);
            case 'z':
                if (env_.dim == 2) return fail("Compute code contains undefinded token 'z'");
                ++pos_;
                return push_ref(2);
            case 'u': ++pos_; return push_ref(16);
            case 'v': ++pos_; return push_ref(17);
            case 'w':
                if (env_.dim == 2) return fail("Compute code contains undefinded token 'w'");
                ++pos_;
                return push_ref(18);
            default: break;
        }
        if (pos_ + 1 < len_ && ((c == 'f' && src_[pos_ + 1] >= '0' && src_[pos_ + 1] <= '7') ||
                                (c == 'c' && src_[pos_ + 1] >= '0' && src_[pos_ + 1] <= '3')))
            return fail(std::string("Identifier '") + c + src_[pos_ + 1] +
                        "' is no longer allowed, see documentation for more information.");
        if (c == '_' || std::isalpha(static_cast<unsigned char>(c)))
        {
            const size_t b = pos_;
            while (pos_ < len_ &&
                   (src_[pos_] == '_' || std::isalnum(static_cast<unsigned char>(src_[pos_]))))
                ++pos_;
            const std::string id(src_ + b, pos_ - b);
            for (uint32_t k = 0; k < env_.nvol; ++k)
                if (env_.vol_names && env_.vol_names[k] && id == env_.vol_names[k]) return push_ref(4 + k);
            for (uint32_t k = 0; k < env_.nconst; ++k)
                if (env_.const_names && env_.const_names[k] && id == env_.const_names[k])
                    return push_ref(12 + k);
            return fail("Invalid identifier '" + id + "'.");
        }
        // Constant <- [-]?[0-9]+[.][0-9]+ / [-]?[0-9]+
        size_t q = pos_;
        if (q < len_ && src_[q] == '-') ++q;
        const size_t d0 = q;
        while (q < len_ && std::isdigit(static_cast<unsigned char>(src_[q]))) ++q;
        if (q > d0)
        {
            if (q + 1 < len_ && src_[q] == '.' && std::isdigit(static_cast<unsigned char>(src_[q + 1])))
            {
                ++q;
                while (q < len_ && std::isdigit(static_cast<unsigned char>(src_[q]))) ++q;
            }
            const std::string tok(src_ + pos_, q - pos_);
            pos_ = q;
            return push_const(std::strtof(tok.c_str(), nullptr));
        }
        return fail(std::string("syntax error, unexpected '") + c + "'");
    }
};

} // namespace

extern "C" void gr_set_error(const char * msg); // gr_capi.cpp

extern "C" int gr_expr_compile(const char * src, const gr_expr_env * env, gr_program * out)
{
    if (!src || !env || !out)
    {
        gr_set_error("gr_expr_compile: null argument");
        return GR_E_INVALID;
    }
    std::memset(out, 0, sizeof(*out));
    Compiler c(src, *env, *out);
    std::string err;
    if (!c.run(err))
    {
        std::string m = err + " (in '" + src + "')";
        gr_set_error(m.c_str());
        std::memset(out, 0, sizeof(*out));
        return GR_E_PARSE;
    }
    return GR_OK;
}

extern "C" int gr_expr_eval_host(const gr_program * pr, const float * ref, float * out)
{
    if (!pr || !ref || !out) return GR_E_INVALID;
    float st[GR_MAX_STACK];
    int sp = 0;
    for (uint32_t i = 0; i < pr->n; ++i)
    {
        const uint32_t w = pr->code[i], op = w & 0xffu, arg = w >> 8;
        switch (op)
        {
            case gr::OP_REF: st[sp++] = ref[arg]; break;
            case gr::OP_CONST: st[sp++] = pr->imm[arg]; break;
            case gr::OP_ADD: sp--; st[sp - 1] = st[sp - 1] + st[sp]; break;
            case gr::OP_SUB: sp--; st[sp - 1] = st[sp - 1] - st[sp]; break;
            case gr::OP_MUL: sp--; st[sp - 1] = st[sp - 1] * st[sp]; break;
            case gr::OP_DIV: sp--; st[sp - 1] = st[sp - 1] / st[sp]; break;
            case gr::OP_SQRT: st[sp - 1] = std::sqrt(st[sp - 1]); break;
            case gr::OP_EXP: st[sp - 1] = std::exp(st[sp - 1]); break;
            case gr::OP_LN: st[sp - 1] = std::log(st[sp - 1]); break;
            case gr::OP_SIN: st[sp - 1] = std::sin(st[sp - 1]); break;
            case gr::OP_COS: st[sp - 1] = std::cos(st[sp - 1]); break;
            case gr::OP_TAN: st[sp - 1] = std::tan(st[sp - 1]); break;
            case gr::OP_ASIN: st[sp - 1] = std::asin(st[sp - 1]); break;
            case gr::OP_ACOS: st[sp - 1] = std::acos(st[sp - 1]); break;
            case gr::OP_ATAN: st[sp - 1] = std::atan(st[sp - 1]); break;
            case gr::OP_ATAN2: sp--; st[sp - 1] = std::atan2(st[sp - 1], st[sp]); break;
            case gr::OP_ABS: st[sp - 1] = std::fabs(st[sp - 1]); break;
            default: return GR_E_INVALID;
        }
    }
    *out = sp > 0 ? st[0] : 0.f;
    return GR_OK;
}

namespace gr
{

static uint32_t dense_ref(uint32_t slot)
{
    if (slot <= 3) return slot;           // x y z cv
    if (slot >= 16) return slot - 16 + 4; // u v w
    return 7 + slot;                      // volumes, constants
}

static uint32_t unary_sub(uint32_t op)
{
    switch (op)
    {
        case OP_SQRT: return 0;
        case OP_EXP: return 1;
        case OP_LN: return 2;
        case OP_SIN: return 3;
        case OP_COS: return 4;
        case OP_TAN: return 5;
        case OP_ASIN: return 6;
        case OP_ACOS: return 7;
        case OP_ATAN: return 8;
        default: return 9; // OP_ABS
    }
}

// Postfix -> device form. A symbolic stack tracks which values are still unemitted leaves
// (reference slots or immediates); a binary operation with a leaf operand becomes one leaf-form
// instruction, and leaves are pushed only when an operation needs them on the stack. The device
// stack then holds exactly the emitted (materialized) values in their symbolic order.
int lower_program(const gr_program & in, Program & out, uint32_t & refmask)
{
    struct Sym
    {
        uint32_t kind; // 0 on the device stack, 1 reference leaf, 2 immediate leaf
        uint32_t arg;  // slot or immediate index
    };
    Sym st[GR_MAX_STACK + 1];
    int sp = 0;
    uint32_t n = 0;
    std::memcpy(out.imm, in.imm, sizeof(out.imm));
    auto emit = [&](uint32_t cls, uint32_t sub, uint32_t arg) -> bool {
        if (n >= GR_MAX_CODE) return false;
        out.code[n++] = (cls << 4 | sub) | (arg << 8);
        return true;
    };
    auto leaf_arg = [&](const Sym & x) -> uint32_t {
        if (x.kind == 1)
        {
            refmask |= 1u << x.arg;
            return dense_ref(x.arg);
        }
        return x.arg;
    };
    auto push = [&](Sym & x) -> bool {
        if (x.kind && !emit(DC_PUSH, x.kind == 2 ? 1u : 0u, leaf_arg(x))) return false;
        x.kind = 0;
        return true;
    };
    if (in.n > GR_MAX_CODE) return GR_E_INVALID;
    // register-slot reference at code[k] (dense index < 7), or -1
    auto reg_ref = [&](uint32_t k) -> int {
        if (k >= in.n || (in.code[k] & 0xffu) != OP_REF) return -1;
        const uint32_t a = in.code[k] >> 8;
        if (a >= 19) return -1;
        const uint32_t d = dense_ref(a);
        return d < 7 ? int(d) : -1;
    };
    auto op_at = [&](uint32_t k) -> uint32_t { return k < in.n ? in.code[k] & 0xffu : 0u; };
    for (uint32_t i = 0; i < in.n; ++i)
    {
        const uint32_t op = in.code[i] & 0xffu, arg = in.code[i] >> 8;
        // a*b + c*d [+ e*f] over register slots: postfix r r * r r * + [r r * +]
        if (op == OP_REF && sp < GR_MAX_STACK && reg_ref(i) >= 0 && reg_ref(i + 1) >= 0 &&
            op_at(i + 2) == OP_MUL && reg_ref(i + 3) >= 0 && reg_ref(i + 4) >= 0 &&
            op_at(i + 5) == OP_MUL && op_at(i + 6) == OP_ADD)
        {
            uint32_t packed = 0, terms = 2;
            const uint32_t at[6] = {i, i + 1, i + 3, i + 4, i + 7, i + 8};
            const bool third = reg_ref(i + 7) >= 0 && reg_ref(i + 8) >= 0 && op_at(i + 9) == OP_MUL &&
                               op_at(i + 10) == OP_ADD;
            if (third) terms = 3;
            for (uint32_t k = 0; k < 2 * terms; ++k)
            {
                packed |= uint32_t(reg_ref(at[k])) << (3 * k);
                refmask |= 1u << (in.code[at[k]] >> 8);
            }
            // a square root of the sum right after it joins the instruction
            const uint32_t last = i + (third ? 10 : 6);
            const bool root = op_at(last + 1) == OP_SQRT;
            if (!emit(DC_DOT, terms | (root ? 4u : 0u), packed)) return GR_E_INVALID;
            st[sp++] = Sym{0, 0};
            i = last + (root ? 1 : 0);
            continue;
        }
        if (op == OP_REF || op == OP_CONST)
        {
            if (sp >= GR_MAX_STACK || (op == OP_REF && arg >= 19) || (op == OP_CONST && arg >= GR_MAX_CODE))
                return GR_E_INVALID;
            st[sp++] = Sym{op == OP_REF ? 1u : 2u, arg};
            continue;
        }
        const bool binary = (op >= OP_ADD && op <= OP_DIV) || op == OP_ATAN2;
        if (binary)
        {
            if (sp < 2) return GR_E_INVALID;
            Sym r = st[--sp], l = st[--sp];
            bool ok = true;
            const uint32_t k = op == OP_ATAN2 ? 4u : op - OP_ADD;
            if (k == 4)
            {
                if (l.kind && !r.kind) ok = push(l) && emit(DC_BINARY, 4u | 8u, 0); // [.., r, l]
                else ok = push(l) && push(r) && emit(DC_BINARY, 4u, 0);
            }
            else if (r.kind) // l is (or becomes) the top: top op leaf
                ok = push(l) && emit(DC_LEAF, k | (r.kind == 2 ? 8u : 0u), leaf_arg(r));
            else if (l.kind) // r is the top: leaf op top
                ok = emit(DC_LEAF, k | 4u | (l.kind == 2 ? 8u : 0u), leaf_arg(l));
            else
                ok = emit(DC_BINARY, k, 0);
            if (!ok) return GR_E_INVALID;
            st[sp++] = Sym{0, 0};
            continue;
        }
        if (op < OP_SQRT || op > OP_ABS || op == OP_ATAN2 || sp < 1) return GR_E_INVALID;
        if (!push(st[sp - 1]) || !emit(DC_UNARY, unary_sub(op), 0)) return GR_E_INVALID;
    }
    if (sp != 1 || !push(st[0])) return GR_E_INVALID;
    out.n = n;
    return GR_OK;
}

static float apply(uint32_t k, float a, float b)
{
    switch (k)
    {
        case 0: return a + b;
        case 1: return a - b;
        case 2: return a * b;
        case 3: return a / b;
        default: return std::atan2(a, b);
    }
}

float eval_lowered(const Program & pr, const float * ref)
{
    static const uint32_t slot_of[7] = {0, 1, 2, 3, 16, 17, 18};
    auto refv = [&](uint32_t d) { return d < 7 ? ref[slot_of[d]] : ref[d - 7]; };
    float acc = 0.f, acc2 = 0.f, deep[GR_MAX_STACK];
    uint32_t depth = 0;
    for (uint32_t i = 0; i < pr.n; ++i)
    {
        const uint32_t w = pr.code[i], cls = (w >> 4) & 0xfu, sub = w & 0xfu, arg = w >> 8;
        if (cls == DC_LEAF)
        {
            const float x = (sub & 8) ? pr.imm[arg] : refv(arg);
            acc = (sub & 4) ? apply(sub & 3, x, acc) : apply(sub & 3, acc, x);
        }
        else if (cls == DC_PUSH || cls == DC_DOT)
        {
            float x;
            if (cls == DC_DOT)
            {
                auto r = [&](uint32_t k) { return refv((arg >> (3 * k)) & 7u); };
                x = (r(0) * r(1)) + (r(2) * r(3));
                if ((sub & 3) == 3) x = x + (r(4) * r(5));
                if (sub & 4) x = std::sqrt(x);
            }
            else
                x = sub ? pr.imm[arg] : refv(arg);
            if (depth >= 2) deep[depth - 2] = acc2;
            acc2 = acc;
            acc = x;
            ++depth;
        }
        else if (cls == DC_BINARY)
        {
            const bool rev = sub & 8;
            acc = rev ? apply(sub & 7, acc, acc2) : apply(sub & 7, acc2, acc);
            --depth;
            if (depth >= 2) acc2 = deep[depth - 2];
        }
        else
        {
            switch (sub)
            {
                case 0: acc = std::sqrt(acc); break;
                case 1: acc = std::exp(acc); break;
                case 2: acc = std::log(acc); break;
                case 3: acc = std::sin(acc); break;
                case 4: acc = std::cos(acc); break;
                case 5: acc = std::tan(acc); break;
                case 6: acc = std::asin(acc); break;
                case 7: acc = std::acos(acc); break;
                case 8: acc = std::atan(acc); break;
                default: acc = std::fabs(acc); break;
            }
        }
    }
    return acc;
}

} // namespace gr

extern "C" int gr_expr_eval_device_form(const gr_program * pr, const float * ref, float * out)
{
    if (!pr || !ref || !out) return GR_E_INVALID;
    gr::Program lo;
    uint32_t mask = 0;
    if (gr::lower_program(*pr, lo, mask) != GR_OK) return GR_E_INVALID;
    *out = gr::eval_lowered(lo, ref);
    return GR_OK;
}
