// WGSL field bodies -> HIP C++.
//
// linon splices a user-written WGSL function body into
//   fn field_function(p_prev: vec3<f32>, p: vec3<f32>, v0: vec3<f32>, v: vec3<f32>, t: f32) -> vec3<f32>
// (src/main_view.rs:852-861) and lets naga compile it.  Here the body is parsed (a subset of
// WGSL: let/var/const, assignments, if/else, for, while, break/continue/return, scalar and
// vec3<f32> arithmetic, swizzles .x/.y/.z/.xyz, calls of the WGSL built-ins and of the shader's
// helper functions) and emitted as the body of a device functor written against the lw::
// vocabulary (device/wgsl_lib.hpp), which hiprtc then compiles into the ray-cast kernel.
// Every float literal becomes an f32 literal, so nothing is silently computed in double, and
// clamp keeps WGSL's (e, low, high) argument order.
#include "wgsl.hpp"

#include <algorithm>
#include <cctype>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace linon {
namespace {

enum class Tok { Ident, Float, Int, Punct, End };

struct Token {
    Tok kind;
    std::string text;
    int line, col;
};

struct TranslateError : std::runtime_error {
    int line, col;
    TranslateError(const std::string& m, int l, int c) : std::runtime_error(m), line(l), col(c) {}
};

class Lexer {
   public:
    explicit Lexer(const std::string& s) : s_(s) {}
    std::vector<Token> run() {
        std::vector<Token> out;
        while (true) {
            skip_space_and_comments();
            if (i_ >= s_.size()) break;
            out.push_back(next());
        }
        out.push_back(Token{Tok::End, "", line_, col_});
        return out;
    }

   private:
    const std::string& s_;
    size_t i_ = 0;
    int line_ = 1, col_ = 1;

    char peek(size_t k = 0) const { return i_ + k < s_.size() ? s_[i_ + k] : '\0'; }
    void adv() {
        if (s_[i_] == '\n') {
            ++line_;
            col_ = 1;
        } else {
            ++col_;
        }
        ++i_;
    }
    void skip_space_and_comments() {
        while (i_ < s_.size()) {
            char c = peek();
            if (std::isspace((unsigned char)c)) {
                adv();
            } else if (c == '/' && peek(1) == '/') {
                while (i_ < s_.size() && peek() != '\n') adv();
            } else if (c == '/' && peek(1) == '*') {
                int depth = 0;  // WGSL block comments nest
                do {
                    if (peek() == '/' && peek(1) == '*') {
                        adv();
                        adv();
                        ++depth;
                    } else if (peek() == '*' && peek(1) == '/') {
                        adv();
                        adv();
                        --depth;
                    } else {
                        if (i_ >= s_.size()) throw TranslateError("unterminated block comment", line_, col_);
                        adv();
                    }
                } while (depth > 0);
            } else if ((unsigned char)c >= 0x80) {
                // non-ASCII outside comments (e.g. the degree sign) is not WGSL
                throw TranslateError("unexpected non-ASCII character", line_, col_);
            } else {
                break;
            }
        }
    }
    Token next() {
        int l = line_, c = col_;
        char ch = peek();
        if (std::isalpha((unsigned char)ch) || ch == '_') {
            std::string id;
            while (std::isalnum((unsigned char)peek()) || peek() == '_') {
                id += peek();
                adv();
            }
            return Token{Tok::Ident, id, l, c};
        }
        if (std::isdigit((unsigned char)ch) || (ch == '.' && std::isdigit((unsigned char)peek(1)))) return number(l, c);
        static const char* ops[] = {"<<=", ">>=", "&&", "||", "==", "!=", "<=", ">=", "+=", "-=", "*=", "/=",
                                    "%=",  "&=",  "|=", "^=", "++", "--", "->", "<<", ">>"};
        for (const char* op : ops) {
            size_t n = std::strlen(op);
            if (s_.compare(i_, n, op) == 0) {
                for (size_t k = 0; k < n; ++k) adv();
                return Token{Tok::Punct, op, l, c};
            }
        }
        if (std::strchr("+-*/%<>=!&|^~()[]{},;:.@", ch)) {
            adv();
            return Token{Tok::Punct, std::string(1, ch), l, c};
        }
        throw TranslateError(std::string("unexpected character '") + ch + "'", l, c);
    }
    Token number(int l, int c) {
        std::string t;
        bool is_float = false;
        if (peek() == '0' && (peek(1) == 'x' || peek(1) == 'X')) {
            t += "0x";
            adv();
            adv();
            while (std::isxdigit((unsigned char)peek()) || peek() == '.') {
                if (peek() == '.') is_float = true;
                t += peek();
                adv();
            }
            if (peek() == 'p' || peek() == 'P') {
                is_float = true;
                t += peek();
                adv();
                if (peek() == '+' || peek() == '-') {
                    t += peek();
                    adv();
                }
                while (std::isdigit((unsigned char)peek())) {
                    t += peek();
                    adv();
                }
            }
        } else {
            while (std::isdigit((unsigned char)peek())) {
                t += peek();
                adv();
            }
            if (peek() == '.') {
                is_float = true;
                t += '.';
                adv();
                while (std::isdigit((unsigned char)peek())) {
                    t += peek();
                    adv();
                }
            }
            if ((peek() == 'e' || peek() == 'E') &&
                (std::isdigit((unsigned char)peek(1)) ||
                 ((peek(1) == '+' || peek(1) == '-') && std::isdigit((unsigned char)peek(2))))) {
                is_float = true;
                t += 'e';
                adv();
                if (peek() == '+' || peek() == '-') {
                    t += peek();
                    adv();
                }
                while (std::isdigit((unsigned char)peek())) {
                    t += peek();
                    adv();
                }
            }
        }
        char suf = peek();
        if (suf == 'f' || suf == 'h') {
            is_float = true;
            adv();
        } else if (suf == 'i' || suf == 'u') {
            if (is_float) throw TranslateError("integer suffix on a float literal", l, c);
            adv();
            if (suf == 'u') t += "u";
        }
        if (is_float) {
            if (t.back() == '.') t += "0";
            if (t[0] == '.') t = "0" + t;
            return Token{Tok::Float, t, l, c};
        }
        return Token{Tok::Int, t, l, c};
    }
};

bool is_cpp_reserved(const std::string& s) {
    static const char* kw[] = {
        "auto", "char", "class", "default", "delete", "double", "float", "int", "long", "namespace", "new",
        "operator", "private", "protected", "public", "register", "short", "signed", "sizeof", "static",
        "struct", "switch", "template", "this", "throw", "try", "catch", "typedef", "typename", "union",
        "unsigned", "using", "virtual", "void", "volatile", "goto", "friend", "inline", "mutable", "explicit",
        "export", "extern", "enum", "case", "do", "nullptr", "decltype", "constexpr", "static_assert",
        "alignas", "alignof", "noexcept", "thread_local", "asm", "wchar_t", "const_cast", "dynamic_cast",
        "reinterpret_cast", "static_cast", "typeid", "and", "or", "not", "xor", "bitand", "bitor", "compl",
        "lw", "lk", "P", "field", "std", "NULL", "INFINITY", "NAN", "__device__", "__global__"};
    for (const char* k : kw)
        if (s == k) return true;
    return s.rfind("__", 0) == 0;
}

// Expression types: f32, i32, u32, bool, vec3<f32>, and WGSL's abstract int / float for
// literals (abstract values take the type their context asks for, as WGSL's conversion rules
// say; an abstract int left alone becomes i32 and an abstract float f32).
enum class Ty { F, I, U, B, V, AI, AF };

struct Ex {
    std::string code;
    Ty ty;
    bool literal = false;  // a bare abstract literal (concretised by rewriting its text)
};

const char* ty_name(Ty t) {
    switch (t) {
        case Ty::F: return "f32";
        case Ty::I: return "i32";
        case Ty::U: return "u32";
        case Ty::B: return "bool";
        case Ty::V: return "vec3<f32>";
        case Ty::AI: return "abstract-int";
        default: return "abstract-float";
    }
}

// Signatures of the callable vocabulary: 'f' = float-generic (f32 or vec3, component-wise),
// 'n' = numeric-generic (ints stay ints), 'F' f32, 'V' vec3.  Result: 'a' = type of the
// generic argument, 'F' f32, 'V' vec3.
struct Sig {
    const char* name;
    const char* params;
    char result;
};
const Sig kSigs[] = {
    {"sin", "f", 'a'}, {"cos", "f", 'a'}, {"tan", "f", 'a'}, {"asin", "f", 'a'}, {"acos", "f", 'a'},
    {"atan", "f", 'a'}, {"sinh", "f", 'a'}, {"cosh", "f", 'a'}, {"tanh", "f", 'a'}, {"exp", "f", 'a'},
    {"exp2", "f", 'a'}, {"log", "f", 'a'}, {"log2", "f", 'a'}, {"sqrt", "f", 'a'}, {"inverseSqrt", "f", 'a'},
    {"floor", "f", 'a'}, {"ceil", "f", 'a'}, {"round", "f", 'a'}, {"trunc", "f", 'a'}, {"fract", "f", 'a'},
    {"saturate", "f", 'a'}, {"radians", "f", 'a'}, {"degrees", "f", 'a'}, {"abs", "n", 'a'},
    {"sign", "n", 'a'}, {"pow", "ff", 'a'}, {"atan2", "FF", 'F'}, {"min", "nn", 'a'}, {"max", "nn", 'a'},
    {"clamp", "nnn", 'a'}, {"mix", "ffF", 'a'}, {"step", "FF", 'F'}, {"smoothstep", "FFF", 'F'},
    {"fma", "FFF", 'F'}, {"length", "f", 'F'}, {"distance", "VV", 'F'}, {"dot", "VV", 'F'},
    {"cross", "VV", 'V'}, {"normalize", "V", 'V'}, {"reflect", "VV", 'V'},
    // helpers of src/main_view.wgsl:78-146
    {"rotateX", "VF", 'V'}, {"rotateY", "VF", 'V'}, {"rotateZ", "VF", 'V'}, {"translate", "VFFF", 'V'},
    {"refraction_index", "F", 'F'}, {"refraction", "FFVV", 'V'}, {"point_plane_distance", "VVV", 'F'},
    {"sigmoid", "F", 'F'},
};

class Parser {
   public:
    explicit Parser(std::vector<Token> toks) : t_(std::move(toks)) {
        scopes_.push_back({{"p_prev", Ty::V}, {"p", Ty::V}, {"v0", Ty::V}, {"v", Ty::V}, {"t", Ty::F},
                           {"PI", Ty::F}});
        scopes_.push_back({});
    }

    std::string body() {
        std::ostringstream o;
        bool returns = false;
        while (!at_end()) {
            o << statement(2);
            returns = last_returns_;
        }
        // WGSL: a function with a return type must end in a statement that always returns
        if (!returns) fail("field_function must end with a return statement (every path must return a vec3<f32>)");
        return o.str();
    }

   private:
    std::vector<Token> t_;
    size_t p_ = 0;
    std::vector<std::vector<std::pair<std::string, Ty>>> scopes_;
    bool last_returns_ = false;  // did the statement just parsed always return?

    // the token stream ends in one End token, which neither cur() nor take() ever moves past
    const Token& cur() const { return t_[std::min(p_, t_.size() - 1)]; }
    const Token& look(size_t k) const { return t_[std::min(p_ + k, t_.size() - 1)]; }
    bool at_end() const { return cur().kind == Tok::End; }
    bool is(const char* s) const { return cur().kind == Tok::Punct && cur().text == s; }
    bool is_kw(const char* s) const { return cur().kind == Tok::Ident && cur().text == s; }
    [[noreturn]] void fail(const std::string& m) const { throw TranslateError(m, cur().line, cur().col); }
    Token take() {
        Token t = cur();
        if (p_ + 1 < t_.size()) ++p_;
        return t;
    }
    void expect(const char* s) {
        if (!is(s)) fail(std::string("expected '") + s + "', found '" + (at_end() ? "end of body" : cur().text) + "'");
        ++p_;
    }
    static std::string cname(const std::string& s) { return is_cpp_reserved(s) ? s + "_w" : s; }
    static std::string ind(int d) { return std::string(4 * d, ' '); }

    void declare(const std::string& n, Ty t) { scopes_.back().push_back({n, t}); }
    bool lookup(const std::string& n, Ty& t) const {
        for (auto s = scopes_.rbegin(); s != scopes_.rend(); ++s)
            for (auto v = s->rbegin(); v != s->rend(); ++v)
                if (v->first == n) {
                    t = v->second;
                    return true;
                }
        return false;
    }

    // ---- concretisation ---------------------------------------------------------------
    static bool floaty(Ty t) { return t == Ty::F || t == Ty::AF; }
    Ex to(const Ex& e, Ty want) const {
        if (e.ty == want) return e;
        if (e.ty == Ty::AI) {
            const bool dec_lit = e.literal && e.code.rfind("0x", 0) != 0;
            if (want == Ty::F) return {dec_lit ? e.code + ".0f" : "float(" + e.code + ")", Ty::F};
            if (want == Ty::I) return {e.code, Ty::I};
            if (want == Ty::U) return {e.literal ? e.code + "u" : "unsigned(" + e.code + ")", Ty::U};
        }
        if (e.ty == Ty::AF && want == Ty::F) return {e.code, Ty::F};
        if (want == Ty::V && (e.ty == Ty::F || e.ty == Ty::AF || e.ty == Ty::AI))
            fail(std::string("a ") + ty_name(e.ty) + " value is used where vec3<f32> is required");
        fail(std::string("type mismatch: ") + ty_name(e.ty) + " where " + ty_name(want) + " is required");
    }
    Ex concrete(const Ex& e) const {
        if (e.ty == Ty::AI) return to(e, Ty::I);
        if (e.ty == Ty::AF) return to(e, Ty::F);
        return e;
    }
    static const char* cty(Ty t) {
        switch (t) {
            case Ty::F: return "float";
            case Ty::I: return "int";
            case Ty::U: return "unsigned";
            case Ty::B: return "bool";
            default: return "vec3";
        }
    }

    // type := f32 | i32 | u32 | bool | vec3<f32> | vec3f
    Ty type() {
        std::string n = take().text;
        if (n == "f32") return Ty::F;
        if (n == "i32") return Ty::I;
        if (n == "u32") return Ty::U;
        if (n == "bool") return Ty::B;
        if (n == "vec3f") return Ty::V;
        if (n == "vec3") {
            expect("<");
            if (take().text != "f32") fail("only vec3<f32> vectors are supported");
            expect(">");
            return Ty::V;
        }
        --p_;
        fail("unsupported type '" + n + "' (supported: f32, i32, u32, bool, vec3<f32>)");
    }

    std::string block(int d) {
        expect("{");
        scopes_.push_back({});
        std::ostringstream o;
        o << "{\n";
        bool ret = false;
        while (!is("}")) {
            if (at_end()) fail("missing '}'");
            o << statement(d + 1);
            ret = ret || last_returns_;
        }
        expect("}");
        scopes_.pop_back();
        o << ind(d) << "}";
        last_returns_ = ret;
        return o.str();
    }

    std::string simple(bool allow_decl) {
        if (allow_decl && (is_kw("let") || is_kw("var") || is_kw("const"))) {
            std::string kw = take().text;
            if (cur().kind != Tok::Ident) fail("expected a name after '" + kw + "'");
            std::string raw = take().text;
            bool typed = false;
            Ty ty = Ty::F;
            if (is(":")) {
                ++p_;
                ty = type();
                typed = true;
            }
            std::string decl;
            if (is("=")) {
                ++p_;
                Ex e = expr();
                Ex c = typed ? to(e, ty) : concrete(e);
                ty = c.ty;
                decl = (kw == "var" ? "" : "const ") + std::string(cty(ty)) + " " + cname(raw) + " = " + c.code;
            } else {
                if (kw != "var") fail("'" + kw + "' needs an initialiser");
                if (!typed) fail("'var' without a type needs an initialiser");
                decl = std::string(cty(ty)) + " " + cname(raw) + "{}";
            }
            declare(raw, ty);
            return decl;
        }
        if (cur().kind == Tok::Ident && cur().text == "_" && look(1).text == "=") {
            p_ += 2;
            return "(void)(" + expr().code + ")";
        }
        Ex lhs = postfix_expr();
        if (is("++") || is("--")) {
            if (lhs.ty != Ty::I && lhs.ty != Ty::U) fail("++/-- need an integer variable");
            return lhs.code + take().text;
        }
        static const char* cops[] = {"+=", "-=", "*=", "/=", "%="};
        for (const char* op : cops) {
            if (is(op)) {
                ++p_;
                Ex rhs = expr();
                Ex r = arith(std::string(op, 1), lhs, rhs);
                return lhs.code + " = " + to(r, lhs.ty).code;
            }
        }
        if (is("=")) {
            ++p_;
            return lhs.code + " = " + to(expr(), lhs.ty).code;
        }
        return "(void)(" + lhs.code + ")";
    }

    std::string statement(int d) {
        last_returns_ = false;
        std::ostringstream o;
        if (is(";")) {
            ++p_;
            return "";
        }
        if (is("{")) {
            o << ind(d) << block(d) << "\n";
            return o.str();
        }
        if (is_kw("return")) {
            last_returns_ = true;
            ++p_;
            if (is(";")) fail("field_function must return a vec3<f32>");
            Ex e = expr();
            if (e.ty != Ty::V) fail(std::string("field_function must return a vec3<f32>, not ") + ty_name(e.ty));
            o << ind(d) << "return " << e.code << ";\n";
            expect(";");
            return o.str();
        }
        if (is_kw("if")) {
            ++p_;
            o << ind(d) << "if (" << cond() << ") " << block(d);
            bool all = last_returns_, has_else = false;
            while (is_kw("else")) {
                ++p_;
                if (is_kw("if")) {
                    ++p_;
                    o << " else if (" << cond() << ") " << block(d);
                    all = all && last_returns_;
                } else {
                    o << " else " << block(d);
                    all = all && last_returns_;
                    has_else = true;
                    break;
                }
            }
            o << "\n";
            last_returns_ = all && has_else;
            return o.str();
        }
        if (is_kw("for")) {
            ++p_;
            expect("(");
            scopes_.push_back({});
            std::string init = is(";") ? "" : simple(true);
            expect(";");
            std::string c = is(";") ? "" : cond();
            expect(";");
            std::string upd = is(")") ? "" : simple(false);
            expect(")");
            o << ind(d) << "for (" << init << "; " << c << "; " << upd << ") " << block(d) << "\n";
            scopes_.pop_back();
            last_returns_ = false;
            return o.str();
        }
        if (is_kw("while")) {
            ++p_;
            o << ind(d) << "while (" << cond() << ") " << block(d) << "\n";
            last_returns_ = false;
            return o.str();
        }
        if (is_kw("loop")) {
            ++p_;
            o << ind(d) << "while (true) " << block(d) << "\n";
            last_returns_ = false;
            return o.str();
        }
        if (is_kw("break") || is_kw("continue")) {
            std::string k = take().text;
            if (k == "break" && is_kw("if")) {
                ++p_;
                o << ind(d) << "if (" << cond() << ") break;\n";
            } else {
                o << ind(d) << k << ";\n";
            }
            expect(";");
            return o.str();
        }
        if (is_kw("discard") || is_kw("switch") || is_kw("fn") || is_kw("struct") || is_kw("continuing"))
            fail("'" + cur().text + "' is not supported in a field function body");
        o << ind(d) << simple(true) << ";\n";
        expect(";");
        return o.str();
    }

    std::string cond() {
        Ex e = expr();
        if (e.ty != Ty::B) fail(std::string("condition must be bool, not ") + ty_name(e.ty));
        return e.code;
    }

    // ---- expressions (WGSL precedence) ----------------------------------------------
    Ex expr() { return binary(1); }

    static int prec(const std::string& op) {
        if (op == "||") return 1;
        if (op == "&&") return 2;
        if (op == "|") return 3;
        if (op == "^") return 4;
        if (op == "&") return 5;
        if (op == "==" || op == "!=") return 6;
        if (op == "<" || op == ">" || op == "<=" || op == ">=") return 7;
        if (op == "<<" || op == ">>") return 8;
        if (op == "+" || op == "-") return 9;
        if (op == "*" || op == "/" || op == "%") return 10;
        return -1;
    }

    // arithmetic typing: vec3 with vec3 or f32 (either side), scalars of one type; abstract
    // operands adopt the other side's type
    Ex arith(const std::string& op, Ex a, Ex b) {
        auto scal = [](Ty t) { return t != Ty::V && t != Ty::B; };
        if (a.ty == Ty::B || b.ty == Ty::B) fail("arithmetic on bool");
        Ty rt;
        if (a.ty == Ty::V || b.ty == Ty::V) {
            if (a.ty != Ty::V) a = to(a, Ty::F);
            if (b.ty != Ty::V) b = to(b, Ty::F);
            rt = Ty::V;
        } else if (a.ty == Ty::AI && b.ty == Ty::AI) {
            rt = Ty::AI;
        } else if (floaty(a.ty) || floaty(b.ty)) {
            if (a.ty == Ty::AF && b.ty == Ty::AF) {
                rt = Ty::AF;
            } else {
                a = to(a, Ty::F);
                b = to(b, Ty::F);
                rt = Ty::F;
            }
        } else {
            if (!scal(a.ty) || !scal(b.ty)) fail("bad operands");
            Ty it = a.ty == Ty::AI ? b.ty : a.ty;
            a = to(a, it);
            b = to(b, it);
            rt = it;
        }
        if (op == "%") return {"lw::wmod(" + a.code + ", " + b.code + ")", rt};
        if (op == "/" && (rt == Ty::I || rt == Ty::U || rt == Ty::AI))  // WGSL's total integer division
            return {"lw::wdiv(" + a.code + ", " + b.code + ")", rt};
        return {"(" + a.code + " " + op + " " + b.code + ")", rt};
    }

    Ex binary(int min_prec) {
        Ex lhs = unary();
        while (cur().kind == Tok::Punct) {
            std::string op = cur().text;
            int pr = prec(op);
            if (pr < 0 || pr < min_prec) break;
            ++p_;
            Ex rhs = binary(pr + 1);
            if (pr >= 9) {
                lhs = arith(op, lhs, rhs);
            } else if (pr >= 6) {  // comparisons
                if (pr != 6 && (lhs.ty == Ty::V || rhs.ty == Ty::V)) fail("vector comparison is not supported");
                if (lhs.ty == Ty::V || rhs.ty == Ty::V) fail("vector comparison is not supported");
                Ex s = arith("+", lhs, rhs);  // unify operand types
                Ty ut = s.ty == Ty::AI ? Ty::I : (s.ty == Ty::AF ? Ty::F : s.ty);
                if (lhs.ty == Ty::B && rhs.ty == Ty::B) ut = Ty::B;
                Ex a = lhs.ty == Ty::B ? lhs : to(lhs, ut), b = rhs.ty == Ty::B ? rhs : to(rhs, ut);
                lhs = {"(" + a.code + " " + op + " " + b.code + ")", Ty::B};
            } else if (op == "&&" || op == "||") {
                if (lhs.ty != Ty::B || rhs.ty != Ty::B) fail("'" + op + "' needs bool operands");
                lhs = {"(" + lhs.code + " " + op + " " + rhs.code + ")", Ty::B};
            } else {
                fail("bitwise operator '" + op + "' is not supported");
            }
        }
        return lhs;
    }

    Ex unary() {
        if (is("-")) {
            ++p_;
            Ex e = unary();
            if (e.ty == Ty::B) fail("'-' on bool");
            return {"(-" + e.code + ")", e.ty, false};
        }
        if (is("!")) {
            ++p_;
            Ex e = unary();
            if (e.ty != Ty::B) fail("'!' needs a bool");
            return {"(!" + e.code + ")", Ty::B};
        }
        if (is("&") || is("*") || is("~")) fail("'" + cur().text + "' is not supported in a field function body");
        return postfix_expr();
    }

    Ex postfix_expr() {
        Ex e = primary();
        while (true) {
            if (is(".")) {
                ++p_;
                if (cur().kind != Tok::Ident) fail("expected a component after '.'");
                if (e.ty != Ty::V) fail("member access on a non-vector value");
                std::string comp = take().text;
                auto idx = [&](char c) -> int {
                    if (c == 'x' || c == 'r') return 0;
                    if (c == 'y' || c == 'g') return 1;
                    if (c == 'z' || c == 'b') return 2;
                    fail("unsupported component '" + std::string(1, c) + "'");
                };
                if (comp.size() == 1) {
                    static const char xyz[] = "xyz";
                    e = {"(" + e.code + ")." + std::string(1, xyz[idx(comp[0])]), Ty::F};
                } else if (comp.size() == 3) {
                    e = {"lw::swizzle3<" + std::to_string(idx(comp[0])) + "," + std::to_string(idx(comp[1])) + "," +
                             std::to_string(idx(comp[2])) + ">(" + e.code + ")",
                         Ty::V};
                } else {
                    fail("only 1- and 3-component swizzles are supported ('." + comp + "')");
                }
            } else if (is("[")) {
                fail("indexing is not supported in a field function body");
            } else {
                break;
            }
        }
        return e;
    }

    std::vector<Ex> args() {
        expect("(");
        std::vector<Ex> a;
        while (!is(")")) {
            a.push_back(expr());
            if (is(",")) {
                ++p_;
                continue;
            }
            if (!is(")")) fail("expected ',' or ')' in an argument list");
        }
        expect(")");
        return a;
    }

    Ex call(const std::string& fname, std::vector<Ex> a) {
        const Sig* sig = nullptr;
        for (const Sig& s : kSigs)
            if (fname == s.name) sig = &s;
        if (!sig) fail("unknown function '" + fname + "'");
        const size_t np = std::strlen(sig->params);
        if (a.size() != np)
            fail("'" + fname + "' takes " + std::to_string(np) + " argument(s), got " + std::to_string(a.size()));
        // the generic type: vec3 if any generic argument is vec3, else f32 if any is float,
        // else the concrete integer type, else abstract int
        Ty g = Ty::AI;
        for (size_t i = 0; i < np; ++i) {
            char k = sig->params[i];
            if (k != 'f' && k != 'n') continue;
            Ty t = a[i].ty;
            if (t == Ty::B) fail("'" + fname + "' does not take bool");
            if (t == Ty::V) g = Ty::V;
            else if (g != Ty::V && floaty(t)) g = Ty::F;
            else if (g == Ty::AI && (t == Ty::I || t == Ty::U)) g = t;
        }
        bool float_only = std::strchr(sig->params, 'f') != nullptr;
        if (g == Ty::AI && float_only) g = Ty::F;
        std::string s = "lw::" + fname + "(";
        for (size_t i = 0; i < np; ++i) {
            char k = sig->params[i];
            Ty want = k == 'F' ? Ty::F : k == 'V' ? Ty::V : g;
            if (want == Ty::AI) want = Ty::I;
            if (k == 'f' && (want == Ty::I || want == Ty::U)) want = Ty::F;
            s += (i ? ", " : "") + to(a[i], want).code;
        }
        s += ")";
        Ty rt = sig->result == 'F' ? Ty::F : sig->result == 'V' ? Ty::V : (g == Ty::AI ? Ty::I : g);
        return {s, rt};
    }

    Ex primary() {
        const Token& tk = cur();
        if (tk.kind == Tok::Float) {
            ++p_;
            return {tk.text + "f", Ty::AF, true};
        }
        if (tk.kind == Tok::Int) {
            ++p_;
            if (!tk.text.empty() && tk.text.back() == 'u') return {tk.text, Ty::U, true};
            return {tk.text, Ty::AI, true};
        }
        if (is("(")) {
            ++p_;
            Ex e = expr();
            expect(")");
            return {"(" + e.code + ")", e.ty, false};
        }
        if (tk.kind != Tok::Ident) fail("unexpected '" + (tk.kind == Tok::End ? std::string("end of body") : tk.text) +
                                        "' in an expression");
        std::string id = tk.text;
        if (id == "true" || id == "false") {
            ++p_;
            return {id, Ty::B};
        }
        if (id == "vec3" || id == "vec3f") {
            ++p_;
            if (id == "vec3" && is("<")) {
                ++p_;
                if (take().text != "f32") fail("only vec3<f32> vectors are supported");
                expect(">");
            }
            std::vector<Ex> a = args();
            if (a.size() == 1 && a[0].ty == Ty::V) return {"lw::vec3(" + a[0].code + ")", Ty::V};
            if (a.size() != 3 && a.size() != 1 && !a.empty()) fail("vec3<f32>(...) takes 0, 1 or 3 arguments");
            std::string s = "lw::vec3f(";
            if (a.empty()) s += "0.0f";
            for (size_t i = 0; i < a.size(); ++i) {
                Ex c = a[i];
                if (c.ty == Ty::I || c.ty == Ty::U) c = {"float(" + c.code + ")", Ty::F};
                s += (i ? ", " : "") + to(c, Ty::F).code;
            }
            return {s + ")", Ty::V};
        }
        if (id == "vec2" || id == "vec4" || id == "vec2f" || id == "vec4f" || id == "mat3x3" || id == "mat2x2" ||
            id == "array")
            fail("type '" + id + "' is not supported in a field function body (use vec3<f32> and scalars)");
        if (id == "f32" || id == "i32" || id == "u32" || id == "bool") {
            ++p_;
            std::vector<Ex> a = args();
            if (a.size() != 1) fail(id + "(...) takes one argument");
            if (a[0].ty == Ty::V) fail(id + "(...) of a vector");
            Ty t = id == "f32" ? Ty::F : id == "i32" ? Ty::I : id == "u32" ? Ty::U : Ty::B;
            if (a[0].ty == Ty::AI || a[0].ty == Ty::AF) {
                if (t == Ty::F) return to(a[0], Ty::F);
                if (a[0].ty == Ty::AI) return to(a[0], t);
            }
            return {std::string(cty(t)) + "(" + a[0].code + ")", t};
        }
        if (id == "select") {
            ++p_;
            std::vector<Ex> a = args();
            if (a.size() != 3) fail("select(f, t, cond) takes three arguments");
            if (a[2].ty != Ty::B) fail("select's third argument must be bool");
            Ex u = (a[0].ty == Ty::V || a[1].ty == Ty::V) ? Ex{"", Ty::V} : arith("+", a[0], a[1]);
            Ty ut = u.ty == Ty::AI ? Ty::I : (u.ty == Ty::AF ? Ty::F : u.ty);
            return {"((" + a[2].code + ") ? (" + to(a[1], ut).code + ") : (" + to(a[0], ut).code + "))", ut};
        }
        ++p_;
        if (is("(")) return call(id, args());
        Ty vt;
        if (!lookup(id, vt)) {
            --p_;
            fail("unknown identifier '" + id + "'");
        }
        return {id == "PI" ? "lw::PI" : cname(id), vt};
    }
};

}  // namespace

bool translate_wgsl_body(const std::string& body, std::string& out, std::string& err) {
    try {
        Lexer lx(body);
        Parser ps(lx.run());
        out = ps.body();
        return true;
    } catch (const TranslateError& e) {
        err = "field_function:" + std::to_string(e.line) + ":" + std::to_string(e.col) + ": " + e.what();
        return false;
    }
}

std::string user_field_source(const std::string& translated) {
    std::ostringstream o;
    o << "namespace lw {\n"
         "struct UserField {\n"
         "    __device__ __forceinline__ vec3 operator()(vec3 p_prev, vec3 p, vec3 v0, vec3 v, float t) const {\n"
         "        (void)p_prev; (void)p; (void)v0; (void)v; (void)t;\n"
         "#line 1 \"field_function\"\n"
      << translated
      << "    }\n"
         "};\n"
         "}  // namespace lw\n"
         "extern \"C\" __global__ void __launch_bounds__(64) LK_RAYCAST_OCCUPANCY linon_raycast_user(RaycastParams P) {\n"
         "    lk::raycast_body<lw::UserField>(P);\n"
         "}\n";
    return o.str();
}

}  // namespace linon
