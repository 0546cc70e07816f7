// lua.cpp -- the subset of Lua 5.1 that D+'s symmetry scripts use, evaluated on
// the host to run a script's Populate(p, nlayers) (LuaSymmetry,
// Backend/Backend/Symmetry.cpp:404-560).  The reference embeds the Lua 5.1
// interpreter; this image has none, so the language the scripts need is
// implemented here from the Lua 5.1 reference manual:
//
//   values     nil, booleans, numbers (double), strings, tables, functions
//   statements local / global assignment (multiple targets), if / elseif /
//              else, numeric and generic for (pairs, ipairs), while, repeat,
//              do ... end, break, return, function definitions (global, local,
//              table fields, methods)
//   expressions arithmetic (+ - * / % ^ unary -), comparison, and / or / not,
//              concatenation (..), length (#), table constructors, calls and
//              method calls, varargs are not supported
//   library    math (pi, huge, sin, cos, tan, asin, acos, atan, atan2, sinh,
//              cosh, tanh, exp, log, log10, sqrt, abs, floor, ceil, fmod, modf,
//              pow, min, max, deg, rad), table.getn / insert, string.format
//              (%d %i %f %g %e %s), tostring, tonumber, type, error, assert,
//              print (to stderr), pairs, ipairs, unpack
//
// As in the reference, os and io are absent.  A script error, a runaway loop
// (more than 2 * 10^7 statements and loop steps) or a missing Populate raises ERROR_INVALIDARGS
// with the script's message.

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "device.hpp"

namespace dplus {
namespace lua {

struct Table;
struct Func;
struct Value {
    enum Kind { Nil, Bool, Num, Str, Tab, Fn } k = Nil;
    bool b = false;
    double n = 0.0;
    std::shared_ptr<std::string> s;
    std::shared_ptr<Table> t;
    std::shared_ptr<Func> f;
    static Value num(double x) {
        Value v;
        v.k = Num;
        v.n = x;
        return v;
    }
    static Value boolean(bool x) {
        Value v;
        v.k = Bool;
        v.b = x;
        return v;
    }
    static Value str(std::string x) {
        Value v;
        v.k = Str;
        v.s = std::make_shared<std::string>(std::move(x));
        return v;
    }
    bool truthy() const { return !(k == Nil || (k == Bool && !b)); }
};
using Values = std::vector<Value>;

[[noreturn]] void fail(const std::string &m) { throw Error(kErrInvalidArgs, "symmetry script: " + m); }

std::string fmt_num(double x) {
    if (x == std::floor(x) && std::fabs(x) < 1e15) {
        char b[64];
        snprintf(b, sizeof b, "%.0f", x);
        return b;
    }
    char b[64];
    snprintf(b, sizeof b, "%.14g", x);
    return b;
}

struct Key {
    int k;  // 0 number, 1 string, 2 boolean, 3 table/function identity
    double n;
    std::string s;
    const void *p;
    bool operator<(const Key &o) const {
        if (k != o.k) return k < o.k;
        if (k == 0 || k == 2) return n < o.n;
        if (k == 1) return s < o.s;
        return p < o.p;
    }
};

struct Table {
    std::vector<Value> arr;  // arr[i] holds key i + 1
    std::map<Key, std::pair<Value, Value>> hash;
    static bool as_index(const Value &k, size_t &i) {
        if (k.k != Value::Num || k.n < 1 || k.n != std::floor(k.n) || k.n > 1e9) return false;
        i = (size_t)k.n - 1;
        return true;
    }
    static Key key_of(const Value &k) {
        Key K{0, 0.0, "", nullptr};
        switch (k.k) {
            case Value::Num: K.k = 0, K.n = k.n; break;
            case Value::Str: K.k = 1, K.s = *k.s; break;
            case Value::Bool: K.k = 2, K.n = k.b; break;
            case Value::Tab: K.k = 3, K.p = k.t.get(); break;
            case Value::Fn: K.k = 3, K.p = k.f.get(); break;
            default: fail("table index is nil");
        }
        return K;
    }
    Value get(const Value &k) const {
        size_t i;
        if (as_index(k, i)) {
            if (i < arr.size()) return arr[i];
        }
        if (k.k == Value::Nil) return Value();
        auto it = hash.find(key_of(k));
        return it == hash.end() ? Value() : it->second.second;
    }
    void set(const Value &k, const Value &v) {
        size_t i;
        if (as_index(k, i)) {
            if (i < arr.size()) {
                arr[i] = v;
                if (v.k == Value::Nil && i + 1 == arr.size()) arr.pop_back();
                return;
            }
            if (i == arr.size() && v.k != Value::Nil) {
                arr.push_back(v);
                // migrate following keys from the hash part
                for (;;) {
                    Value nk = Value::num((double)arr.size() + 1);
                    auto it = hash.find(key_of(nk));
                    if (it == hash.end()) break;
                    arr.push_back(it->second.second);
                    hash.erase(it);
                }
                return;
            }
        }
        if (k.k == Value::Nil) fail("table index is nil");
        if (v.k == Value::Nil)
            hash.erase(key_of(k));
        else
            hash[key_of(k)] = {k, v};
    }
    double length() const { return (double)arr.size(); }
};

// ------------------------------------------------------------------ lexer
enum Tok { TEnd, TName, TNum, TStr, TOp, TKw };
struct Token {
    Tok t;
    std::string s;
    double n = 0.0;
    int line = 0;
};

static const char *kKeywords[] = {"and", "break", "do", "else", "elseif", "end", "false", "for", "function", "if",
                                  "in", "local", "nil", "not", "or", "repeat", "return", "then", "true", "until",
                                  "while"};

std::vector<Token> lex(const std::string &src) {
    std::vector<Token> out;
    size_t i = 0;
    int line = 1;
    const size_t n = src.size();
    auto long_bracket = [&](size_t at, size_t &endpos, std::string &body) -> bool {
        // [[...]] or [=[...]=]
        if (at >= n || src[at] != '[') return false;
        size_t j = at + 1;
        int eq = 0;
        while (j < n && src[j] == '=') eq++, j++;
        if (j >= n || src[j] != '[') return false;
        const std::string close = "]" + std::string(eq, '=') + "]";
        const size_t e = src.find(close, j + 1);
        if (e == std::string::npos) fail("unfinished long string");
        body = src.substr(j + 1, e - j - 1);
        for (char c : body)
            if (c == '\n') line++;
        endpos = e + close.size();
        return true;
    };
    while (i < n) {
        const char c = src[i];
        if (c == '\n') {
            line++;
            i++;
            continue;
        }
        if (isspace((unsigned char)c)) {
            i++;
            continue;
        }
        if (c == '-' && i + 1 < n && src[i + 1] == '-') {
            size_t e;
            std::string body;
            if (long_bracket(i + 2, e, body)) {
                i = e;
                continue;
            }
            while (i < n && src[i] != '\n') i++;
            continue;
        }
        Token tk;
        tk.line = line;
        if (isalpha((unsigned char)c) || c == '_') {
            size_t j = i;
            while (j < n && (isalnum((unsigned char)src[j]) || src[j] == '_')) j++;
            tk.s = src.substr(i, j - i);
            tk.t = TName;
            for (const char *k : kKeywords)
                if (tk.s == k) tk.t = TKw;
            out.push_back(tk);
            i = j;
            continue;
        }
        if (isdigit((unsigned char)c) || (c == '.' && i + 1 < n && isdigit((unsigned char)src[i + 1]))) {
            char *end = nullptr;
            if (c == '0' && i + 1 < n && (src[i + 1] == 'x' || src[i + 1] == 'X'))
                tk.n = (double)strtoll(src.c_str() + i, &end, 16);
            else
                tk.n = strtod(src.c_str() + i, &end);
            tk.t = TNum;
            out.push_back(tk);
            i = (size_t)(end - src.c_str());
            continue;
        }
        if (c == '"' || c == '\'') {
            std::string s;
            size_t j = i + 1;
            while (j < n && src[j] != c) {
                if (src[j] == '\n') fail("unfinished string at line " + std::to_string(line));
                if (src[j] == '\\' && j + 1 < n) {
                    const char e = src[++j];
                    switch (e) {
                        case 'n': s += '\n'; break;
                        case 't': s += '\t'; break;
                        case 'r': s += '\r'; break;
                        case '\\': s += '\\'; break;
                        case '"': s += '"'; break;
                        case '\'': s += '\''; break;
                        case '\n': s += '\n'; line++; break;
                        default:
                            if (isdigit((unsigned char)e)) {
                                int v = 0, k = 0;
                                while (k < 3 && j < n && isdigit((unsigned char)src[j])) v = 10 * v + (src[j++] - '0'), k++;
                                j--;
                                s += (char)v;
                            } else {
                                s += e;
                            }
                    }
                    j++;
                    continue;
                }
                s += src[j++];
            }
            if (j >= n) fail("unfinished string at line " + std::to_string(line));
            tk.t = TStr;
            tk.s = s;
            out.push_back(tk);
            i = j + 1;
            continue;
        }
        {
            size_t e;
            std::string body;
            if (c == '[' && long_bracket(i, e, body)) {
                tk.t = TStr;
                tk.s = body.size() && body[0] == '\n' ? body.substr(1) : body;
                out.push_back(tk);
                i = e;
                continue;
            }
        }
        static const char *ops3[] = {"..."};
        static const char *ops2[] = {"==", "~=", "<=", ">=", ".."};
        tk.t = TOp;
        bool got = false;
        for (const char *o : ops3)
            if (src.compare(i, 3, o) == 0) tk.s = o, i += 3, got = true;
        if (!got)
            for (const char *o : ops2)
                if (!got && src.compare(i, 2, o) == 0) tk.s = o, i += 2, got = true;
        if (!got) {
            if (!strchr("+-*/%^#<>=(){}[];:,.", c)) fail(std::string("unexpected symbol '") + c + "' at line " +
                                                         std::to_string(line));
            tk.s = std::string(1, c);
            i++;
        }
        out.push_back(tk);
    }
    Token e;
    e.t = TEnd;
    e.line = line;
    out.push_back(e);
    return out;
}

// ------------------------------------------------------------------ syntax tree
struct Expr;
struct Stat;
using ExprP = std::shared_ptr<Expr>;
using StatP = std::shared_ptr<Stat>;
using Block = std::vector<StatP>;

struct FuncDef {
    std::vector<std::string> params;
    bool is_method = false;
    std::shared_ptr<Block> body;
};

struct Expr {
    enum K { Nil, True, False, Number, String, Name, Index, Call, Method, Function, Binop, Unop, TableCons, Paren } k;
    double n = 0.0;
    std::string s;  // name, string, operator, method name
    ExprP a, b;     // operands / object and key / callee
    std::vector<ExprP> args;
    std::vector<std::pair<ExprP, ExprP>> fields;  // table constructor: key (null: positional), value
    std::shared_ptr<FuncDef> fn;
    int line = 0;
};

struct Stat {
    enum K { Assign, Local, CallStat, If, While, Repeat, NumFor, GenFor, Do, Return, Break, LocalFunction } k;
    std::vector<ExprP> targets, exprs;
    std::vector<std::string> names;
    std::vector<ExprP> conds;
    std::vector<std::shared_ptr<Block>> blocks;  // if: one per cond (+ else), loops: body
    std::shared_ptr<FuncDef> fn;
    int line = 0;
};

struct Parser {
    std::vector<Token> tk;
    size_t p = 0;
    const Token &cur() const { return tk[p]; }
    bool is_op(const char *o) const { return cur().t == TOp && cur().s == o; }
    bool is_kw(const char *k) const { return cur().t == TKw && cur().s == k; }
    [[noreturn]] void err(const std::string &m) const { fail(m + " near line " + std::to_string(cur().line)); }
    void expect_op(const char *o) {
        if (!is_op(o)) err(std::string("'") + o + "' expected");
        p++;
    }
    void expect_kw(const char *k) {
        if (!is_kw(k)) err(std::string("'") + k + "' expected");
        p++;
    }
    std::string name() {
        if (cur().t != TName) err("name expected");
        return tk[p++].s;
    }

    bool block_end() const {
        return cur().t == TEnd || is_kw("end") || is_kw("else") || is_kw("elseif") || is_kw("until");
    }

    std::shared_ptr<Block> block() {
        auto b = std::make_shared<Block>();
        while (!block_end()) {
            if (is_op(";")) {
                p++;
                continue;
            }
            StatP s = statement();
            b->push_back(s);
            if (s->k == Stat::Return || s->k == Stat::Break) {
                if (is_op(";")) p++;
                break;
            }
        }
        return b;
    }

    std::shared_ptr<FuncDef> funcbody(bool method) {
        auto f = std::make_shared<FuncDef>();
        f->is_method = method;
        if (method) f->params.push_back("self");
        expect_op("(");
        if (!is_op(")")) {
            for (;;) {
                if (is_op("...")) err("varargs are not supported");
                f->params.push_back(name());
                if (!is_op(",")) break;
                p++;
            }
        }
        expect_op(")");
        f->body = block();
        expect_kw("end");
        return f;
    }

    StatP statement() {
        auto s = std::make_shared<Stat>();
        s->line = cur().line;
        if (is_kw("if")) {
            p++;
            s->k = Stat::If;
            s->conds.push_back(expr());
            expect_kw("then");
            s->blocks.push_back(block());
            while (is_kw("elseif")) {
                p++;
                s->conds.push_back(expr());
                expect_kw("then");
                s->blocks.push_back(block());
            }
            if (is_kw("else")) {
                p++;
                s->blocks.push_back(block());
            }
            expect_kw("end");
            return s;
        }
        if (is_kw("while")) {
            p++;
            s->k = Stat::While;
            s->conds.push_back(expr());
            expect_kw("do");
            s->blocks.push_back(block());
            expect_kw("end");
            return s;
        }
        if (is_kw("repeat")) {
            p++;
            s->k = Stat::Repeat;
            s->blocks.push_back(block());
            expect_kw("until");
            s->conds.push_back(expr());
            return s;
        }
        if (is_kw("do")) {
            p++;
            s->k = Stat::Do;
            s->blocks.push_back(block());
            expect_kw("end");
            return s;
        }
        if (is_kw("for")) {
            p++;
            const std::string n1 = name();
            if (is_op("=")) {
                p++;
                s->k = Stat::NumFor;
                s->names.push_back(n1);
                s->exprs.push_back(expr());
                expect_op(",");
                s->exprs.push_back(expr());
                if (is_op(",")) {
                    p++;
                    s->exprs.push_back(expr());
                }
            } else {
                s->k = Stat::GenFor;
                s->names.push_back(n1);
                while (is_op(",")) {
                    p++;
                    s->names.push_back(name());
                }
                expect_kw("in");
                s->exprs = explist();
            }
            expect_kw("do");
            s->blocks.push_back(block());
            expect_kw("end");
            return s;
        }
        if (is_kw("function")) {
            p++;
            // funcname: Name {'.' Name} [':' Name]
            auto target = std::make_shared<Expr>();
            target->k = Expr::Name;
            target->s = name();
            bool method = false;
            while (is_op(".") || is_op(":")) {
                const bool colon = is_op(":");
                p++;
                auto key = std::make_shared<Expr>();
                key->k = Expr::String;
                key->s = name();
                auto ix = std::make_shared<Expr>();
                ix->k = Expr::Index;
                ix->a = target;
                ix->b = key;
                target = ix;
                if (colon) {
                    method = true;
                    break;
                }
            }
            s->k = Stat::Assign;
            s->targets.push_back(target);
            auto fe = std::make_shared<Expr>();
            fe->k = Expr::Function;
            fe->fn = funcbody(method);
            s->exprs.push_back(fe);
            return s;
        }
        if (is_kw("local")) {
            p++;
            if (is_kw("function")) {
                p++;
                s->k = Stat::LocalFunction;
                s->names.push_back(name());
                s->fn = funcbody(false);
                return s;
            }
            s->k = Stat::Local;
            s->names.push_back(name());
            while (is_op(",")) {
                p++;
                s->names.push_back(name());
            }
            if (is_op("=")) {
                p++;
                s->exprs = explist();
            }
            return s;
        }
        if (is_kw("return")) {
            p++;
            s->k = Stat::Return;
            if (!block_end() && !is_op(";")) s->exprs = explist();
            return s;
        }
        if (is_kw("break")) {
            p++;
            s->k = Stat::Break;
            return s;
        }
        // assignment or call
        ExprP e = suffixed();
        if (is_op("=") || is_op(",")) {
            s->k = Stat::Assign;
            s->targets.push_back(e);
            while (is_op(",")) {
                p++;
                s->targets.push_back(suffixed());
            }
            expect_op("=");
            s->exprs = explist();
            for (auto &t : s->targets)
                if (t->k != Expr::Name && t->k != Expr::Index) err("cannot assign to this expression");
            return s;
        }
        if (e->k != Expr::Call && e->k != Expr::Method) err("syntax error");
        s->k = Stat::CallStat;
        s->exprs.push_back(e);
        return s;
    }

    std::vector<ExprP> explist() {
        std::vector<ExprP> v{expr()};
        while (is_op(",")) {
            p++;
            v.push_back(expr());
        }
        return v;
    }

    ExprP primary() {
        auto e = std::make_shared<Expr>();
        e->line = cur().line;
        if (cur().t == TName) {
            e->k = Expr::Name;
            e->s = tk[p++].s;
            return e;
        }
        if (is_op("(")) {
            p++;
            e->k = Expr::Paren;
            e->a = expr();
            expect_op(")");
            return e;
        }
        err("unexpected symbol");
    }

    std::vector<ExprP> call_args() {
        std::vector<ExprP> a;
        if (cur().t == TStr) {
            auto e = std::make_shared<Expr>();
            e->k = Expr::String;
            e->s = tk[p++].s;
            a.push_back(e);
            return a;
        }
        if (is_op("{")) {
            a.push_back(table_cons());
            return a;
        }
        expect_op("(");
        if (!is_op(")")) a = explist();
        expect_op(")");
        return a;
    }

    ExprP suffixed() {
        ExprP e = primary();
        for (;;) {
            if (is_op(".")) {
                p++;
                auto ix = std::make_shared<Expr>();
                ix->k = Expr::Index;
                ix->a = e;
                auto key = std::make_shared<Expr>();
                key->k = Expr::String;
                key->s = name();
                ix->b = key;
                e = ix;
            } else if (is_op("[")) {
                p++;
                auto ix = std::make_shared<Expr>();
                ix->k = Expr::Index;
                ix->a = e;
                ix->b = expr();
                expect_op("]");
                e = ix;
            } else if (is_op(":")) {
                p++;
                auto m = std::make_shared<Expr>();
                m->k = Expr::Method;
                m->line = cur().line;
                m->a = e;
                m->s = name();
                m->args = call_args();
                e = m;
            } else if (is_op("(") || is_op("{") || cur().t == TStr) {
                auto c = std::make_shared<Expr>();
                c->k = Expr::Call;
                c->line = cur().line;
                c->a = e;
                c->args = call_args();
                e = c;
            } else {
                return e;
            }
        }
    }

    ExprP table_cons() {
        auto t = std::make_shared<Expr>();
        t->k = Expr::TableCons;
        expect_op("{");
        while (!is_op("}")) {
            if (is_op("[")) {
                p++;
                ExprP k = expr();
                expect_op("]");
                expect_op("=");
                t->fields.push_back({k, expr()});
            } else if (cur().t == TName && tk[p + 1].t == TOp && tk[p + 1].s == "=") {
                auto k = std::make_shared<Expr>();
                k->k = Expr::String;
                k->s = tk[p].s;
                p += 2;
                t->fields.push_back({k, expr()});
            } else {
                t->fields.push_back({nullptr, expr()});
            }
            if (is_op(",") || is_op(";"))
                p++;
            else
                break;
        }
        expect_op("}");
        return t;
    }

    ExprP simple() {
        auto e = std::make_shared<Expr>();
        e->line = cur().line;
        if (cur().t == TNum) {
            e->k = Expr::Number;
            e->n = tk[p++].n;
            return e;
        }
        if (cur().t == TStr) {
            e->k = Expr::String;
            e->s = tk[p++].s;
            return e;
        }
        if (is_kw("nil")) return p++, e->k = Expr::Nil, e;
        if (is_kw("true")) return p++, e->k = Expr::True, e;
        if (is_kw("false")) return p++, e->k = Expr::False, e;
        if (is_op("{")) return table_cons();
        if (is_kw("function")) {
            p++;
            e->k = Expr::Function;
            e->fn = funcbody(false);
            return e;
        }
        if (is_op("...")) err("varargs are not supported");
        return suffixed();
    }

    static int left_prec(const std::string &o) {
        if (o == "or") return 1;
        if (o == "and") return 2;
        if (o == "<" || o == ">" || o == "<=" || o == ">=" || o == "~=" || o == "==") return 3;
        if (o == "..") return 5;  // right associative (right prec 4)
        if (o == "+" || o == "-") return 6;
        if (o == "*" || o == "/" || o == "%") return 7;
        if (o == "^") return 10;  // right associative (right prec 9)
        return -1;
    }
    static int right_prec(const std::string &o) {
        if (o == "..") return 4;
        if (o == "^") return 9;
        return left_prec(o);
    }
    static constexpr int kUnaryPrec = 8;

    std::string binop() const {
        if (cur().t == TOp || (cur().t == TKw && (cur().s == "and" || cur().s == "or"))) return cur().s;
        return "";
    }

    ExprP expr(int limit = 0) {
        ExprP e;
        if (is_kw("not") || is_op("-") || is_op("#")) {
            auto u = std::make_shared<Expr>();
            u->k = Expr::Unop;
            u->line = cur().line;
            u->s = tk[p++].s;
            u->a = expr(kUnaryPrec);
            e = u;
        } else {
            e = simple();
        }
        for (;;) {
            const std::string o = binop();
            const int lp = o.empty() ? -1 : left_prec(o);
            if (lp < 0 || lp <= limit) break;
            p++;
            auto b = std::make_shared<Expr>();
            b->k = Expr::Binop;
            b->line = cur().line;
            b->s = o;
            b->a = e;
            b->b = expr(right_prec(o));
            e = b;
        }
        return e;
    }
};

// ------------------------------------------------------------------ evaluation
struct Scope {
    std::unordered_map<std::string, Value> vars;
    std::shared_ptr<Scope> parent;
};
using ScopeP = std::shared_ptr<Scope>;

using Builtin = std::function<Values(Values &)>;
struct Func {
    std::shared_ptr<FuncDef> def;
    ScopeP closure;
    Builtin builtin;
};

struct BreakSignal {};
struct ReturnSignal {
    Values v;
};

constexpr long long kMaxSteps = 20000000LL;  // about a second of interpretation

struct Interp {
    std::shared_ptr<Table> globals = std::make_shared<Table>();
    long long steps = 0;

    Value gget(const std::string &n) { return globals->get(Value::str(n)); }
    void gset(const std::string &n, const Value &v) { globals->set(Value::str(n), v); }

    Value *lookup(const ScopeP &sc, const std::string &n) {
        for (Scope *s = sc.get(); s; s = s->parent.get()) {
            auto it = s->vars.find(n);
            if (it != s->vars.end()) return &it->second;
        }
        return nullptr;
    }

    static double tonum(const Value &v, int line, const char *what) {
        if (v.k == Value::Num) return v.n;
        if (v.k == Value::Str) {
            char *e = nullptr;
            const double d = strtod(v.s->c_str(), &e);
            if (e && *e == 0 && !v.s->empty()) return d;
        }
        fail(std::string("attempt to perform arithmetic on a ") + type_name(v) + " value (" + what + ") at line " +
             std::to_string(line));
    }
    static const char *type_name(const Value &v) {
        switch (v.k) {
            case Value::Nil: return "nil";
            case Value::Bool: return "boolean";
            case Value::Num: return "number";
            case Value::Str: return "string";
            case Value::Tab: return "table";
            default: return "function";
        }
    }
    static std::string tostr(const Value &v) {
        switch (v.k) {
            case Value::Nil: return "nil";
            case Value::Bool: return v.b ? "true" : "false";
            case Value::Num: return fmt_num(v.n);
            case Value::Str: return *v.s;
            case Value::Tab: return "table";
            default: return "function";
        }
    }
    static bool equal(const Value &a, const Value &b) {
        if (a.k != b.k) return false;
        switch (a.k) {
            case Value::Nil: return true;
            case Value::Bool: return a.b == b.b;
            case Value::Num: return a.n == b.n;
            case Value::Str: return *a.s == *b.s;
            case Value::Tab: return a.t == b.t;
            default: return a.f == b.f;
        }
    }

    Values call(const Value &fv, Values args, int line) {
        if (fv.k != Value::Fn) fail(std::string("attempt to call a ") + type_name(fv) + " value at line " +
                                    std::to_string(line));
        Func &f = *fv.f;
        if (f.builtin) return f.builtin(args);
        auto sc = std::make_shared<Scope>();
        sc->parent = f.closure;
        for (size_t i = 0; i < f.def->params.size(); i++) sc->vars[f.def->params[i]] = i < args.size() ? args[i] : Value();
        try {
            exec_block(*f.def->body, sc);
        } catch (ReturnSignal &r) {
            return std::move(r.v);
        }
        return {};
    }

    Values eval_multi(const ExprP &e, const ScopeP &sc) {
        if (e->k == Expr::Call || e->k == Expr::Method) return eval_call(e, sc);
        return {eval(e, sc)};
    }

    // explist: every expression gives one value except the last, which gives all of its values
    Values eval_list(const std::vector<ExprP> &es, const ScopeP &sc) {
        Values out;
        for (size_t i = 0; i < es.size(); i++) {
            if (i + 1 == es.size()) {
                Values v = eval_multi(es[i], sc);
                out.insert(out.end(), v.begin(), v.end());
            } else {
                out.push_back(eval(es[i], sc));
            }
        }
        return out;
    }

    Values eval_call(const ExprP &e, const ScopeP &sc) {
        if (e->k == Expr::Method) {
            Value obj = eval(e->a, sc);
            Value fn = index(obj, Value::str(e->s), e->line);
            Values args{obj};
            Values rest = eval_list(e->args, sc);
            args.insert(args.end(), rest.begin(), rest.end());
            return call(fn, std::move(args), e->line);
        }
        Value fn = eval(e->a, sc);
        return call(fn, eval_list(e->args, sc), e->line);
    }

    Value index(const Value &obj, const Value &key, int line) {
        if (obj.k == Value::Tab) return obj.t->get(key);
        if (obj.k == Value::Str) {  // string methods through the string library
            Value lib = gget("string");
            if (lib.k == Value::Tab) return lib.t->get(key);
        }
        fail(std::string("attempt to index a ") + type_name(obj) + " value at line " + std::to_string(line));
    }

    Value eval(const ExprP &e, const ScopeP &sc) {
        switch (e->k) {
            case Expr::Nil: return Value();
            case Expr::True: return Value::boolean(true);
            case Expr::False: return Value::boolean(false);
            case Expr::Number: return Value::num(e->n);
            case Expr::String: return Value::str(e->s);
            case Expr::Paren: return eval(e->a, sc);
            case Expr::Name: {
                if (Value *v = lookup(sc, e->s)) return *v;
                return gget(e->s);
            }
            case Expr::Index: return index(eval(e->a, sc), eval(e->b, sc), e->line);
            case Expr::Call:
            case Expr::Method: {
                Values v = eval_call(e, sc);
                return v.empty() ? Value() : v[0];
            }
            case Expr::Function: {
                Value v;
                v.k = Value::Fn;
                v.f = std::make_shared<Func>();
                v.f->def = e->fn;
                v.f->closure = sc;
                return v;
            }
            case Expr::TableCons: {
                Value v;
                v.k = Value::Tab;
                v.t = std::make_shared<Table>();
                double pos = 1;
                for (size_t i = 0; i < e->fields.size(); i++) {
                    const auto &f = e->fields[i];
                    if (f.first) {
                        v.t->set(eval(f.first, sc), eval(f.second, sc));
                    } else if (i + 1 == e->fields.size()) {
                        for (const Value &x : eval_multi(f.second, sc)) v.t->set(Value::num(pos++), x);
                    } else {
                        v.t->set(Value::num(pos++), eval(f.second, sc));
                    }
                }
                return v;
            }
            case Expr::Unop: {
                Value a = eval(e->a, sc);
                if (e->s == "not") return Value::boolean(!a.truthy());
                if (e->s == "-") return Value::num(-tonum(a, e->line, "unary minus"));
                if (a.k == Value::Str) return Value::num((double)a.s->size());
                if (a.k == Value::Tab) return Value::num(a.t->length());
                fail(std::string("attempt to get length of a ") + type_name(a) + " value at line " +
                     std::to_string(e->line));
            }
            case Expr::Binop: {
                const std::string &o = e->s;
                if (o == "and") {
                    Value a = eval(e->a, sc);
                    return a.truthy() ? eval(e->b, sc) : a;
                }
                if (o == "or") {
                    Value a = eval(e->a, sc);
                    return a.truthy() ? a : eval(e->b, sc);
                }
                Value a = eval(e->a, sc), b = eval(e->b, sc);
                if (o == "==") return Value::boolean(equal(a, b));
                if (o == "~=") return Value::boolean(!equal(a, b));
                if (o == "..") {
                    auto cs = [&](const Value &v) {
                        if (v.k != Value::Str && v.k != Value::Num)
                            fail(std::string("attempt to concatenate a ") + type_name(v) + " value at line " +
                                 std::to_string(e->line));
                        return tostr(v);
                    };
                    return Value::str(cs(a) + cs(b));
                }
                if (o == "<" || o == ">" || o == "<=" || o == ">=") {
                    bool r;
                    if (a.k == Value::Num && b.k == Value::Num) {
                        r = o == "<" ? a.n < b.n : o == ">" ? a.n > b.n : o == "<=" ? a.n <= b.n : a.n >= b.n;
                    } else if (a.k == Value::Str && b.k == Value::Str) {
                        const int c = a.s->compare(*b.s);
                        r = o == "<" ? c < 0 : o == ">" ? c > 0 : o == "<=" ? c <= 0 : c >= 0;
                    } else {
                        fail(std::string("attempt to compare ") + type_name(a) + " with " + type_name(b) +
                             " at line " + std::to_string(e->line));
                    }
                    return Value::boolean(r);
                }
                const double x = tonum(a, e->line, o.c_str()), y = tonum(b, e->line, o.c_str());
                if (o == "+") return Value::num(x + y);
                if (o == "-") return Value::num(x - y);
                if (o == "*") return Value::num(x * y);
                if (o == "/") return Value::num(x / y);
                if (o == "%") return Value::num(x - std::floor(x / y) * y);
                if (o == "^") return Value::num(std::pow(x, y));
                fail("unknown operator " + o);
            }
        }
        fail("bad expression");
    }

    void assign(const ExprP &t, const Value &v, const ScopeP &sc) {
        if (t->k == Expr::Name) {
            if (Value *slot = lookup(sc, t->s))
                *slot = v;
            else
                gset(t->s, v);
            return;
        }
        Value obj = eval(t->a, sc);
        if (obj.k != Value::Tab)
            fail(std::string("attempt to index a ") + type_name(obj) + " value at line " + std::to_string(t->line));
        obj.t->set(eval(t->b, sc), v);
    }

    void tick(int line) {
        if (++steps > kMaxSteps) fail("the script ran more than 2 * 10^7 statements (line " + std::to_string(line) + ")");
    }

    void exec_block(const Block &b, const ScopeP &parent) {
        auto sc = std::make_shared<Scope>();
        sc->parent = parent;
        for (const StatP &s : b) exec(*s, sc);
    }

    void exec(const Stat &s, const ScopeP &sc) {
        tick(s.line);
        switch (s.k) {
            case Stat::Local: {
                Values v = s.exprs.empty() ? Values() : eval_list(s.exprs, sc);
                for (size_t i = 0; i < s.names.size(); i++) sc->vars[s.names[i]] = i < v.size() ? v[i] : Value();
                return;
            }
            case Stat::LocalFunction: {
                sc->vars[s.names[0]] = Value();
                Value f;
                f.k = Value::Fn;
                f.f = std::make_shared<Func>();
                f.f->def = s.fn;
                f.f->closure = sc;
                sc->vars[s.names[0]] = f;
                return;
            }
            case Stat::Assign: {
                Values v = eval_list(s.exprs, sc);
                for (size_t i = 0; i < s.targets.size(); i++) assign(s.targets[i], i < v.size() ? v[i] : Value(), sc);
                return;
            }
            case Stat::CallStat: eval_call(s.exprs[0], sc); return;
            case Stat::Do: exec_block(*s.blocks[0], sc); return;
            case Stat::If: {
                for (size_t i = 0; i < s.conds.size(); i++)
                    if (eval(s.conds[i], sc).truthy()) {
                        exec_block(*s.blocks[i], sc);
                        return;
                    }
                if (s.blocks.size() > s.conds.size()) exec_block(*s.blocks.back(), sc);
                return;
            }
            case Stat::While:
                try {
                    while (eval(s.conds[0], sc).truthy()) {
                        tick(s.line);
                        exec_block(*s.blocks[0], sc);
                    }
                } catch (BreakSignal &) {
                }
                return;
            case Stat::Repeat:
                try {
                    for (;;) {
                        tick(s.line);
                        auto inner = std::make_shared<Scope>();
                        inner->parent = sc;
                        for (const StatP &st : *s.blocks[0]) exec(*st, inner);
                        if (eval(s.conds[0], inner).truthy()) break;
                    }
                } catch (BreakSignal &) {
                }
                return;
            case Stat::NumFor: {
                const double a = tonum(eval(s.exprs[0], sc), s.line, "'for' initial value");
                const double b = tonum(eval(s.exprs[1], sc), s.line, "'for' limit");
                const double st = s.exprs.size() > 2 ? tonum(eval(s.exprs[2], sc), s.line, "'for' step") : 1.0;
                if (st == 0) fail("'for' step is zero");
                try {
                    for (double v = a; st > 0 ? v <= b : v >= b; v += st) {
                        tick(s.line);
                        auto inner = std::make_shared<Scope>();
                        inner->parent = sc;
                        inner->vars[s.names[0]] = Value::num(v);
                        for (const StatP &x : *s.blocks[0]) exec(*x, inner);
                    }
                } catch (BreakSignal &) {
                }
                return;
            }
            case Stat::GenFor: {
                Values init = eval_list(s.exprs, sc);
                init.resize(3);
                const Value fn = init[0], state = init[1];
                Value ctl = init[2];
                try {
                    for (;;) {
                        tick(s.line);
                        Values r = call(fn, {state, ctl}, s.line);
                        if (r.empty() || r[0].k == Value::Nil) break;
                        ctl = r[0];
                        auto inner = std::make_shared<Scope>();
                        inner->parent = sc;
                        for (size_t i = 0; i < s.names.size(); i++)
                            inner->vars[s.names[i]] = i < r.size() ? r[i] : Value();
                        for (const StatP &x : *s.blocks[0]) exec(*x, inner);
                    }
                } catch (BreakSignal &) {
                }
                return;
            }
            case Stat::Return: throw ReturnSignal{s.exprs.empty() ? Values() : eval_list(s.exprs, sc)};
            case Stat::Break: throw BreakSignal{};
        }
    }

    // ---------------------------------------------------------------- library
    static Value builtin(Builtin f) {
        Value v;
        v.k = Value::Fn;
        v.f = std::make_shared<Func>();
        v.f->builtin = std::move(f);
        return v;
    }
    static double arg_num(Values &a, size_t i, const char *fn) {
        if (i >= a.size() || (a[i].k != Value::Num && a[i].k != Value::Str))
            fail(std::string("bad argument #") + std::to_string(i + 1) + " to '" + fn + "' (number expected)");
        return tonum(a[i], 0, fn);
    }
    static Value new_table() {
        Value v;
        v.k = Value::Tab;
        v.t = std::make_shared<Table>();
        return v;
    }

    void open_libs() {
        Value math = new_table();
        auto fn1 = [&](const char *n, double (*f)(double)) {
            math.t->set(Value::str(n), builtin([f, n](Values &a) { return Values{Value::num(f(arg_num(a, 0, n)))}; }));
        };
        fn1("sin", std::sin), fn1("cos", std::cos), fn1("tan", std::tan), fn1("asin", std::asin);
        fn1("acos", std::acos), fn1("sinh", std::sinh), fn1("cosh", std::cosh), fn1("tanh", std::tanh);
        fn1("exp", std::exp), fn1("log10", std::log10), fn1("sqrt", std::sqrt), fn1("abs", std::fabs);
        fn1("floor", std::floor), fn1("ceil", std::ceil);
        math.t->set(Value::str("log"), builtin([](Values &a) {
                        const double x = arg_num(a, 0, "log");
                        return Values{Value::num(a.size() > 1 ? std::log(x) / std::log(arg_num(a, 1, "log")) : std::log(x))};
                    }));
        math.t->set(Value::str("atan"), builtin([](Values &a) {
                        return Values{Value::num(a.size() > 1 ? std::atan2(arg_num(a, 0, "atan"), arg_num(a, 1, "atan"))
                                                              : std::atan(arg_num(a, 0, "atan")))};
                    }));
        math.t->set(Value::str("atan2"), builtin([](Values &a) {
                        return Values{Value::num(std::atan2(arg_num(a, 0, "atan2"), arg_num(a, 1, "atan2")))};
                    }));
        math.t->set(Value::str("fmod"), builtin([](Values &a) {
                        return Values{Value::num(std::fmod(arg_num(a, 0, "fmod"), arg_num(a, 1, "fmod")))};
                    }));
        math.t->set(Value::str("pow"), builtin([](Values &a) {
                        return Values{Value::num(std::pow(arg_num(a, 0, "pow"), arg_num(a, 1, "pow")))};
                    }));
        math.t->set(Value::str("modf"), builtin([](Values &a) {
                        double ip;
                        const double fr = std::modf(arg_num(a, 0, "modf"), &ip);
                        return Values{Value::num(ip), Value::num(fr)};
                    }));
        math.t->set(Value::str("deg"), builtin([](Values &a) {
                        return Values{Value::num(arg_num(a, 0, "deg") * 180.0 / kPi)};
                    }));
        math.t->set(Value::str("rad"), builtin([](Values &a) {
                        return Values{Value::num(arg_num(a, 0, "rad") * kPi / 180.0)};
                    }));
        math.t->set(Value::str("min"), builtin([](Values &a) {
                        double m = arg_num(a, 0, "min");
                        for (size_t i = 1; i < a.size(); i++) m = std::min(m, arg_num(a, i, "min"));
                        return Values{Value::num(m)};
                    }));
        math.t->set(Value::str("max"), builtin([](Values &a) {
                        double m = arg_num(a, 0, "max");
                        for (size_t i = 1; i < a.size(); i++) m = std::max(m, arg_num(a, i, "max"));
                        return Values{Value::num(m)};
                    }));
        math.t->set(Value::str("pi"), Value::num(kPi));
        math.t->set(Value::str("huge"), Value::num(HUGE_VAL));
        gset("math", math);

        Value table = new_table();
        table.t->set(Value::str("getn"), builtin([](Values &a) {
                         if (a.empty() || a[0].k != Value::Tab) fail("bad argument #1 to 'getn' (table expected)");
                         return Values{Value::num(a[0].t->length())};
                     }));
        table.t->set(Value::str("insert"), builtin([](Values &a) {
                         if (a.empty() || a[0].k != Value::Tab) fail("bad argument #1 to 'insert' (table expected)");
                         Table &t = *a[0].t;
                         if (a.size() == 2) {
                             t.set(Value::num(t.length() + 1), a[1]);
                         } else if (a.size() >= 3) {
                             const double pos = arg_num(a, 1, "insert");
                             for (double i = t.length(); i >= pos; i--) t.set(Value::num(i + 1), t.get(Value::num(i)));
                             t.set(Value::num(pos), a[2]);
                         }
                         return Values{};
                     }));
        gset("table", table);

        Value string = new_table();
        string.t->set(Value::str("format"), builtin([](Values &a) {
                          if (a.empty() || a[0].k != Value::Str) fail("bad argument #1 to 'format' (string expected)");
                          const std::string &f = *a[0].s;
                          std::string out;
                          size_t ai = 1;
                          for (size_t i = 0; i < f.size(); i++) {
                              if (f[i] != '%') {
                                  out += f[i];
                                  continue;
                              }
                              size_t j = i + 1;
                              while (j < f.size() && strchr("-+ #0123456789.", f[j])) j++;
                              if (j >= f.size()) fail("invalid format string");
                              const char c = f[j];
                              std::string spec = f.substr(i, j - i + 1);
                              char buf[128];
                              if (c == '%') {
                                  out += '%';
                              } else if (c == 'd' || c == 'i') {
                                  spec.back() = 'l';
                                  spec += "ld";
                                  snprintf(buf, sizeof buf, spec.c_str(), (long long)arg_num(a, ai++, "format"));
                                  out += buf;
                              } else if (strchr("fgeEG", c)) {
                                  snprintf(buf, sizeof buf, spec.c_str(), arg_num(a, ai++, "format"));
                                  out += buf;
                              } else if (c == 's') {
                                  out += ai < a.size() ? tostr(a[ai]) : "nil";
                                  ai++;
                              } else {
                                  fail(std::string("unsupported format '%") + c + "'");
                              }
                              i = j;
                          }
                          return Values{Value::str(out)};
                      }));
        gset("string", string);

        gset("tostring", builtin([](Values &a) { return Values{Value::str(a.empty() ? "nil" : tostr(a[0]))}; }));
        gset("tonumber", builtin([](Values &a) {
                 if (a.empty()) return Values{Value()};
                 if (a[0].k == Value::Num) return Values{a[0]};
                 if (a[0].k == Value::Str) {
                     char *e = nullptr;
                     const double d = strtod(a[0].s->c_str(), &e);
                     if (e && *e == 0 && !a[0].s->empty()) return Values{Value::num(d)};
                 }
                 return Values{Value()};
             }));
        gset("type", builtin([](Values &a) { return Values{Value::str(a.empty() ? "nil" : type_name(a[0]))}; }));
        gset("error", builtin([](Values &a) -> Values { fail(a.empty() ? "error" : tostr(a[0])); }));
        gset("assert", builtin([](Values &a) {
                 if (a.empty() || !a[0].truthy()) fail(a.size() > 1 ? tostr(a[1]) : "assertion failed!");
                 return a;
             }));
        gset("print", builtin([](Values &a) {
                 std::string s;
                 for (size_t i = 0; i < a.size(); i++) s += (i ? "\t" : "") + tostr(a[i]);
                 fprintf(stderr, "%s\n", s.c_str());
                 return Values{};
             }));
        gset("unpack", builtin([](Values &a) {
                 if (a.empty() || a[0].k != Value::Tab) fail("bad argument #1 to 'unpack' (table expected)");
                 return a[0].t->arr;
             }));
        Value inext = builtin([](Values &a) {
            const double i = arg_num(a, 1, "ipairs") + 1;
            Value v = a[0].t->get(Value::num(i));
            if (v.k == Value::Nil) return Values{Value()};
            return Values{Value::num(i), v};
        });
        gset("ipairs", builtin([inext](Values &a) {
                 if (a.empty() || a[0].k != Value::Tab) fail("bad argument #1 to 'ipairs' (table expected)");
                 return Values{inext, a[0], Value::num(0)};
             }));
        gset("pairs", builtin([](Values &a) {
                 if (a.empty() || a[0].k != Value::Tab) fail("bad argument #1 to 'pairs' (table expected)");
                 // a snapshot of the keys in array then hash order
                 auto keys = std::make_shared<std::vector<std::pair<Value, Value>>>();
                 const Table &t = *a[0].t;
                 for (size_t i = 0; i < t.arr.size(); i++)
                     if (t.arr[i].k != Value::Nil) keys->push_back({Value::num((double)i + 1), t.arr[i]});
                 for (auto &kv : t.hash) keys->push_back(kv.second);
                 auto pos = std::make_shared<size_t>(0);
                 Value it = builtin([keys, pos](Values &) {
                     if (*pos >= keys->size()) return Values{Value()};
                     const auto &kv = (*keys)[(*pos)++];
                     return Values{kv.first, kv.second};
                 });
                 return Values{it, a[0], Value()};
             }));
    }
};

}  // namespace lua

// Populate(p, nlayers) of a symmetry script (LuaSymmetry::PreCalculate,
// Symmetry.cpp:463-521): p[i][j] (1-based) = params[i - 1][j - 1] for j up to the
// script's Information.NLP (SetTableFromParamVector :607-641; entries a layer
// lacks read as 0), nlayers = params.size(); the result's rows {x, y, z, alpha,
// beta, gamma} (nm, degrees; missing or non-numeric entries read as 0, getlocrot
// :437-461).  nlp: receives Information.NLP.
std::vector<std::array<double, 6>> lua_symmetry_rows(const std::string &script,
                                                     const std::vector<std::vector<double>> &params, int *nlp) {
    using namespace lua;
    Interp I;
    I.open_libs();
    Parser P;
    P.tk = lex(script);
    std::shared_ptr<Block> chunk = P.block();
    if (P.cur().t != TEnd) P.err("'<eof>' expected");
    auto top = std::make_shared<Scope>();
    try {
        I.exec_block(*chunk, top);
    } catch (ReturnSignal &) {
    } catch (BreakSignal &) {
        fail("break outside a loop");
    }
    // Information table (LuaSymmetry::VerifyAndObtainData, :597-630)
    Value info = I.gget("Information");
    if (info.k != Value::Tab) fail("no Information table");
    Value type = info.t->get(Value::str("Type"));
    if (type.k != Value::Str || *type.s != "Symmetry") fail("Information.Type is not \"Symmetry\"");
    Value vn = info.t->get(Value::str("NLP"));
    if (vn.k != Value::Num || vn.n <= 0) fail("Information.NLP must be a positive number");
    const int width = (int)vn.n;
    if (nlp) *nlp = width;
    Value pop = I.gget("Populate");
    if (pop.k != Value::Fn) fail("no Populate function");
    Value p = Interp::new_table();
    for (size_t i = 0; i < params.size(); i++) {
        Value row = Interp::new_table();
        for (int j = 0; j < width; j++)
            row.t->set(Value::num(j + 1.0), Value::num((size_t)j < params[i].size() ? params[i][j] : 0.0));
        p.t->set(Value::num((double)i + 1), row);
    }
    Values r = I.call(pop, {p, Value::num((double)params.size())}, 0);
    if (r.empty() || r[0].k != Value::Tab) fail("Populate did not return a table");
    std::vector<std::array<double, 6>> rows;
    const Table &t = *r[0].t;
    for (double i = 1; i <= t.length(); i++) {
        std::array<double, 6> row{0, 0, 0, 0, 0, 0};
        Value rv = t.get(Value::num(i));
        if (rv.k == Value::Tab)
            for (int c = 0; c < 6; c++) {
                Value x = rv.t->get(Value::num(c + 1.0));
                if (x.k == Value::Num) row[c] = x.n;
            }
        rows.push_back(row);
    }
    return rows;
}

}  // namespace dplus

using namespace dplus;

extern "C" int dplus_lua_symmetry_rows(const char *script, int nlayers, int nlp, const double *params, int capacity,
                                       int *nrows, double *rows) {
    try {
        if (!script || nlayers < 0 || nlp < 0 || (nlayers * nlp > 0 && !params) || !nrows || capacity < 0 ||
            (capacity > 0 && !rows))
            throw Error(kErrInvalidArgs, "bad script arguments");
        std::vector<std::vector<double>> p((size_t)nlayers, std::vector<double>((size_t)nlp));
        for (int i = 0; i < nlayers; i++)
            for (int j = 0; j < nlp; j++) p[i][j] = params[(size_t)i * nlp + j];
        const auto r = lua_symmetry_rows(script, p, nullptr);
        *nrows = (int)r.size();
        if (*nrows > capacity) return kOk;  // the caller retries with capacity >= *nrows
        for (size_t i = 0; i < r.size(); i++)
            for (int c = 0; c < 6; c++) rows[6 * i + c] = r[i][c];
        return kOk;
    } catch (const Error &e) {
        return set_error(e.code, e.what());
    } catch (const std::exception &e) {
        return set_error(kErrGeneral, e.what());
    }
}
