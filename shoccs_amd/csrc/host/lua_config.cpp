// lua_config.cpp -- the reference's Lua `simulation` configuration, read by a
// Lua-subset evaluator written for this library (no Lua interpreter is
// linked), mapped onto shoccs_problem_desc and shoccs_run_desc.
//
// The subset: statements `[local] name = exp` and `name.field[...] = exp`;
// expressions with numbers, strings, booleans, nil, table constructors
// ({a, b}, {k = v}, {[e] = v}), + - * / // % ^ .. unary - not #, comparisons
// and `and`/`or`, parentheses, field and index access, and the math library
// (sqrt, sin, cos, tan, asin, acos, atan, exp, log, abs, floor, ceil, min,
// max, fmod, pi, huge); comments `--` and `--[[ ]]`.  Functions
// (`function (params) body end`, and the statements `[local] function name
// (params) body end`) are values that can be called: their bodies run
// `local` and assignment statements, `if ... then ... elseif ... else ...
// end` and `return explist` (a call in an expression gives its first
// value).  A function sees the locals in scope where it was defined, as they
// were then, and the globals.
//
// The keys read follow the reference's from_lua readers (file:line at each).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "shoccs.h"

namespace {

// ------------------------------------------------------------------- values
struct Table;
struct FuncDef;
struct Value {
    enum Kind { Nil, Bool, Num, Str, Tab, Fun, Builtin } k = Nil;
    double n = 0.0;
    bool b = false;
    std::string s;
    std::shared_ptr<Table> t;
    std::shared_ptr<FuncDef> f;
    static Value num(double v)
    {
        Value x;
        x.k = Num;
        x.n = v;
        return x;
    }
    static Value str(std::string v)
    {
        Value x;
        x.k = Str;
        x.s = std::move(v);
        return x;
    }
    static Value boolean(bool v)
    {
        Value x;
        x.k = Bool;
        x.b = v;
        return x;
    }
    bool truthy() const { return !(k == Nil || (k == Bool && !b)); }
};

struct Table {
    std::map<double, Value> num;
    std::map<std::string, Value> str;
    Value get(const Value& key) const
    {
        if (key.k == Value::Num) {
            auto it = num.find(key.n);
            return it == num.end() ? Value{} : it->second;
        }
        if (key.k == Value::Str) {
            auto it = str.find(key.s);
            return it == str.end() ? Value{} : it->second;
        }
        return Value{};
    }
    Value get(const std::string& k) const { return get(Value::str(k)); }
    Value get(int i) const { return get(Value::num(i)); }
    void set(const Value& key, Value v)
    {
        if (key.k == Value::Num) num[key.n] = std::move(v);
        else if (key.k == Value::Str) str[key.s] = std::move(v);
        else throw std::runtime_error("table index is nil or not a number/string");
    }
    long length() const
    {
        long n = 0;
        while (num.count((double)(n + 1)) && num.at((double)(n + 1)).k != Value::Nil) ++n;
        return n;
    }
};

// ------------------------------------------------------------------- lexer
enum Tok { TEOF, TName, TNum, TStr, TOp };

struct Token {
    Tok t = TEOF;
    std::string s;
    double n = 0.0;
    int line = 1;
};

class Lexer {
  public:
    explicit Lexer(const std::string& src) : src_(&src) { next(); }
    const Token& peek() const { return cur_; }
    Token take()
    {
        Token t = cur_;
        next();
        return t;
    }

  private:
    const std::string* src_;  // the chunk; outlives the lexer
    size_t i_ = 0;
    int line_ = 1;
    Token cur_;

    [[noreturn]] void fail(const std::string& m) const
    {
        throw std::runtime_error("line " + std::to_string(line_) + ": " + m);
    }
    // length of a long bracket opener "[" "="* "[" at i, or 0
    size_t long_open(size_t i, int& level) const
    {
        if (i >= src_->size() || (*src_)[i] != '[') return 0;
        size_t j = i + 1;
        level = 0;
        while (j < src_->size() && (*src_)[j] == '=') ++level, ++j;
        return j < src_->size() && (*src_)[j] == '[' ? j + 1 - i : 0;
    }
    std::string long_body(size_t start, int level)
    {
        const std::string close = "]" + std::string(level, '=') + "]";
        const size_t e = src_->find(close, start);
        if (e == std::string::npos) fail("unfinished long bracket");
        std::string body = src_->substr(start, e - start);
        for (char c : body) line_ += c == '\n';
        i_ = e + close.size();
        return body;
    }
    void skip()
    {
        for (;;) {
            while (i_ < src_->size() && std::isspace((unsigned char)(*src_)[i_])) line_ += (*src_)[i_++] == '\n';
            if (src_->compare(i_, 2, "--") != 0) return;
            i_ += 2;
            int level;
            if (size_t L = long_open(i_, level)) {
                long_body(i_ + L, level);
                continue;
            }
            while (i_ < src_->size() && (*src_)[i_] != '\n') ++i_;
        }
    }
    void next()
    {
        skip();
        cur_ = Token{};
        cur_.line = line_;
        if (i_ >= src_->size()) return;
        const char c = (*src_)[i_];
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t j = i_;
            while (j < src_->size() && (std::isalnum((unsigned char)(*src_)[j]) || (*src_)[j] == '_')) ++j;
            cur_.t = TName;
            cur_.s = src_->substr(i_, j - i_);
            i_ = j;
            return;
        }
        if (std::isdigit((unsigned char)c) || (c == '.' && i_ + 1 < src_->size() && std::isdigit((unsigned char)(*src_)[i_ + 1]))) {
            const char* b = src_->c_str() + i_;
            char* e = nullptr;
            double v;
            if (c == '0' && i_ + 1 < src_->size() && ((*src_)[i_ + 1] == 'x' || (*src_)[i_ + 1] == 'X'))
                v = (double)std::strtoull(b, &e, 16);
            else
                v = std::strtod(b, &e);
            if (e == b) fail("malformed number");
            cur_.t = TNum;
            cur_.n = v;
            i_ += (size_t)(e - b);
            return;
        }
        if (c == '"' || c == '\'') {
            std::string out;
            size_t j = i_ + 1;
            while (j < src_->size() && (*src_)[j] != c) {
                char ch = (*src_)[j++];
                if (ch == '\n') fail("unfinished string");
                if (ch == '\\' && j < src_->size()) {
                    const char e = (*src_)[j++];
                    ch = e == 'n' ? '\n' : e == 't' ? '\t' : e;
                }
                out.push_back(ch);
            }
            if (j >= src_->size()) fail("unfinished string");
            cur_.t = TStr;
            cur_.s = out;
            i_ = j + 1;
            return;
        }
        int level;
        if (size_t L = long_open(i_, level)) {
            cur_.t = TStr;
            cur_.s = long_body(i_ + L, level);
            return;
        }
        static const char* ops[] = {"...", "..", "==", "~=", "<=", ">=", "//", "::"};
        for (const char* op : ops)
            if (src_->compare(i_, std::strlen(op), op) == 0) {
                cur_.t = TOp;
                cur_.s = op;
                i_ += std::strlen(op);
                return;
            }
        if (std::strchr("+-*/%^#&~|<>=(){}[];:,.", c)) {
            cur_.t = TOp;
            cur_.s = std::string(1, c);
            ++i_;
            return;
        }
        fail(std::string("unexpected character '") + c + "'");
    }
};

// A function value: its parameters, the lexer at the first token of its body,
// and the locals in scope where it was defined.
struct FuncDef {
    std::vector<std::string> params;
    Lexer body;
    std::map<std::string, Value> upvalues;
};

// --------------------------------------------------------------- evaluator
class Eval {
  public:
    explicit Eval(const std::string& src) : lx_(src)
    {
        auto m = std::make_shared<Table>();
        const char* fns[] = {"sqrt", "sin", "cos", "tan", "asin", "acos", "atan", "exp",
                             "log", "abs", "floor", "ceil", "min", "max", "fmod"};
        for (const char* f : fns) {
            Value v;
            v.k = Value::Builtin;
            v.s = f;
            m->set(Value::str(f), v);
        }
        m->set(Value::str("pi"), Value::num(3.14159265358979323846));
        m->set(Value::str("huge"), Value::num(std::numeric_limits<double>::infinity()));
        Value mv;
        mv.k = Value::Tab;
        mv.t = m;
        globals_["math"] = mv;
    }

    void run()
    {
        while (lx_.peek().t != TEOF) statement();
    }
    Value global(const std::string& name) const
    {
        auto it = globals_.find(name);
        return it == globals_.end() ? Value{} : it->second;
    }

    // Calls a function value (or a math builtin) with args; its return values.
    std::vector<Value> invoke(const Value& fv, const std::vector<Value>& args)
    {
        if (fv.k == Value::Builtin) return {call(fv, args)};
        if (fv.k != Value::Fun || !fv.f) fail("attempt to call a non-function value");
        if (depth_ >= 200) fail("call depth above 200");
        // the caller's lexer, scope and return state come back on every exit
        struct Restore {
            Eval& e;
            Lexer lx;
            std::map<std::string, Value> locals;
            std::vector<Value> ret;
            bool returned, in_fn;
            ~Restore()
            {
                e.lx_ = lx;
                e.locals_ = std::move(locals);
                e.ret_ = std::move(ret);
                e.returned_ = returned;
                e.in_fn_ = in_fn;
                --e.depth_;
            }
        } guard{*this, lx_, std::move(locals_), std::move(ret_), returned_, in_fn_};
        ++depth_;
        const FuncDef& fd = *fv.f;
        locals_ = fd.upvalues;
        for (size_t i = 0; i < fd.params.size(); ++i) locals_[fd.params[i]] = i < args.size() ? args[i] : Value{};
        lx_ = fd.body;
        ret_.clear();
        returned_ = false;
        in_fn_ = true;
        block();
        if (!returned_ && !is_kw("end")) fail("'end' expected");
        return returned_ ? ret_ : std::vector<Value>{};
    }

  private:
    Lexer lx_;
    std::map<std::string, Value> globals_;
    std::map<std::string, Value> locals_;  // the chunk's, or the running function's, scope
    std::vector<Value> ret_;               // values of the return statement that ran
    bool returned_ = false;                // a return statement ran in the current call
    bool in_fn_ = false;                   // running a function body
    int depth_ = 0;

    [[noreturn]] void fail(const std::string& m) const
    {
        throw std::runtime_error("line " + std::to_string(lx_.peek().line) + ": " + m);
    }
    bool is_op(const char* op) const { return lx_.peek().t == TOp && lx_.peek().s == op; }
    bool is_kw(const char* kw) const { return lx_.peek().t == TName && lx_.peek().s == kw; }
    void expect(const char* op)
    {
        if (!is_op(op)) fail(std::string("'") + op + "' expected");
        lx_.take();
    }
    std::string name()
    {
        if (lx_.peek().t != TName) fail("name expected");
        return lx_.take().s;
    }

    Value lookup(const std::string& n) const
    {
        auto it = locals_.find(n);
        if (it != locals_.end()) return it->second;
        return global(n);
    }

    // assignment target: a name, or a table slot
    struct Target {
        std::string name;
        std::shared_ptr<Table> t;
        Value key;
    };

    // statements up to the end of the block (`end`, `else`, `elseif`, end of
    // the chunk) or a return statement
    void block()
    {
        while (!returned_ && lx_.peek().t != TEOF && !is_kw("end") && !is_kw("else") && !is_kw("elseif"))
            statement();
    }

    static bool opens(const Token& t)
    {
        return t.t == TName && (t.s == "if" || t.s == "function" || t.s == "do" || t.s == "repeat");
    }
    static bool closes(const Token& t) { return t.t == TName && (t.s == "end" || t.s == "until"); }

    // skips a block not taken, to its `else`, `elseif` or `end` (not consumed)
    void skip_block()
    {
        int depth = 0;
        for (;;) {
            const Token& t = lx_.peek();
            if (t.t == TEOF) fail("'end' expected");
            if (depth == 0 && (is_kw("end") || is_kw("else") || is_kw("elseif"))) return;
            if (opens(t)) ++depth;
            else if (closes(t)) --depth;
            lx_.take();
        }
    }

    void expect_kw(const char* kw)
    {
        if (!is_kw(kw)) fail(std::string("'") + kw + "' expected");
        lx_.take();
    }

    // if exp then block {elseif exp then block} [else block] end
    void if_statement()
    {
        bool done = false;
        for (;;) {
            // at `if` or `elseif`
            lx_.take();
            if (done) {
                while (!is_kw("then")) {
                    if (lx_.peek().t == TEOF) fail("'then' expected");
                    lx_.take();
                }
                lx_.take();
                skip_block();
            } else {
                const bool c = expr().truthy();
                expect_kw("then");
                if (c) {
                    block();
                    if (returned_) return;
                    done = true;
                } else {
                    skip_block();
                }
            }
            if (is_kw("elseif")) continue;
            if (is_kw("else")) {
                lx_.take();
                if (done) {
                    skip_block();
                } else {
                    block();
                    if (returned_) return;
                }
            }
            expect_kw("end");
            return;
        }
    }

    void statement()
    {
        if (is_op(";")) {
            lx_.take();
            return;
        }
        if (is_kw("return")) {
            lx_.take();
            if (!in_fn_) fail("return outside a function");
            ret_.clear();
            if (lx_.peek().t != TEOF && !is_kw("end") && !is_kw("else") && !is_kw("elseif") && !is_op(";"))
                ret_ = explist();
            returned_ = true;
            return;
        }
        if (is_kw("if")) {
            if_statement();
            return;
        }
        if (is_kw("function")) {  // function name (params) body end
            lx_.take();
            const std::string n = name();
            Value v = function_value();
            if (locals_.count(n)) locals_[n] = v;
            else globals_[n] = v;
            return;
        }
        if (is_kw("local")) {
            lx_.take();
            if (is_kw("function")) {  // local function name (params) body end
                lx_.take();
                const std::string n = name();
                locals_[n] = Value{};
                locals_[n] = function_value();
                return;
            }
            std::vector<std::string> names{name()};
            while (is_op(",")) {
                lx_.take();
                names.push_back(name());
            }
            std::vector<Value> vals;
            if (is_op("=")) {
                lx_.take();
                vals = explist();
            }
            for (size_t i = 0; i < names.size(); ++i) locals_[names[i]] = i < vals.size() ? vals[i] : Value{};
            return;
        }
        // a call statement (name, fields, then a call) is evaluated for its effects
        {
            Lexer save = lx_;
            name();
            while (is_op(".") || is_op("[")) {
                if (is_op(".")) {
                    lx_.take();
                    name();
                } else {
                    int depth = 0;
                    do {
                        if (lx_.peek().t == TEOF) fail("']' expected");
                        if (is_op("[")) ++depth;
                        else if (is_op("]")) --depth;
                        lx_.take();
                    } while (depth > 0);
                }
            }
            const bool is_call = is_op("(");
            lx_ = save;
            if (is_call) {
                primary();
                return;
            }
        }
        std::vector<Target> targets{target()};
        while (is_op(",")) {
            lx_.take();
            targets.push_back(target());
        }
        expect("=");
        std::vector<Value> vals = explist();
        for (size_t i = 0; i < targets.size(); ++i) {
            Value v = i < vals.size() ? vals[i] : Value{};
            Target& tg = targets[i];
            if (tg.t) tg.t->set(tg.key, v);
            else if (locals_.count(tg.name)) locals_[tg.name] = v;
            else globals_[tg.name] = v;
        }
    }

    Target target()
    {
        Target tg;
        tg.name = name();
        if (!is_op(".") && !is_op("[")) return tg;
        Value cur = lookup(tg.name);
        for (;;) {
            Value key;
            if (is_op(".")) {
                lx_.take();
                key = Value::str(name());
            } else if (is_op("[")) {
                lx_.take();
                key = expr();
                expect("]");
            } else {
                break;
            }
            if (cur.k != Value::Tab) fail("indexing a non-table value");
            if (!is_op(".") && !is_op("[")) {
                tg.t = cur.t;
                tg.key = key;
                return tg;
            }
            cur = cur.t->get(key);
        }
        fail("assignment target expected");
    }

    std::vector<Value> explist()
    {
        std::vector<Value> v{expr()};
        while (is_op(",")) {
            lx_.take();
            v.push_back(expr());
        }
        return v;
    }

    // precedence climbing over Lua's binary operators
    static int prec_left(const std::string& op)
    {
        if (op == "or") return 1;
        if (op == "and") return 2;
        if (op == "<" || op == ">" || op == "<=" || op == ">=" || op == "~=" || op == "==") return 3;
        if (op == "..") return 9;
        if (op == "+" || op == "-") return 10;
        if (op == "*" || op == "/" || op == "//" || op == "%") return 11;
        if (op == "^") return 14;
        return 0;
    }
    std::string binop() const
    {
        const Token& t = lx_.peek();
        if (t.t == TOp || (t.t == TName && (t.s == "and" || t.s == "or"))) return prec_left(t.s) ? t.s : "";
        return "";
    }

    Value expr(int limit = 0)
    {
        Value lhs;
        if (is_op("-") || is_kw("not") || is_op("#")) {
            const std::string op = lx_.take().s;
            Value v = expr(12);  // unary operators bind tighter than all but ^
            if (op == "-") lhs = Value::num(-number(v));
            else if (op == "not") lhs = Value::boolean(!v.truthy());
            else lhs = Value::num(v.k == Value::Tab ? (double)v.t->length() : (double)v.s.size());
        } else {
            lhs = primary();
        }
        for (;;) {
            const std::string op = binop();
            if (op.empty()) break;
            const int p = prec_left(op);
            if (p <= limit) break;
            lx_.take();
            // .. and ^ are right associative
            const int rlimit = (op == ".." || op == "^") ? p - 1 : p;
            if (op == "and") {
                Value r = expr(rlimit);
                lhs = lhs.truthy() ? r : lhs;
            } else if (op == "or") {
                Value r = expr(rlimit);
                lhs = lhs.truthy() ? lhs : r;
            } else {
                lhs = arith(op, lhs, expr(rlimit));
            }
        }
        return lhs;
    }

    double number(const Value& v) const
    {
        if (v.k == Value::Num) return v.n;
        if (v.k == Value::Str) {
            char* e = nullptr;
            const double d = std::strtod(v.s.c_str(), &e);
            if (e && *e == 0 && !v.s.empty()) return d;
        }
        fail("arithmetic on a non-number value");
    }

    static std::string tostr(const Value& v)
    {
        if (v.k == Value::Str) return v.s;
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.14g", v.n);
        return buf;
    }

    Value arith(const std::string& op, const Value& a, const Value& b) const
    {
        if (op == "..") return Value::str(tostr(a) + tostr(b));
        if (op == "==" || op == "~=") {
            bool eq = a.k == b.k && ((a.k == Value::Num && a.n == b.n) || (a.k == Value::Str && a.s == b.s) ||
                                     (a.k == Value::Bool && a.b == b.b) || a.k == Value::Nil ||
                                     (a.k == Value::Tab && a.t == b.t));
            return Value::boolean(op == "==" ? eq : !eq);
        }
        if (op == "<" || op == ">" || op == "<=" || op == ">=") {
            bool r;
            if (a.k == Value::Str && b.k == Value::Str)
                r = op == "<" ? a.s < b.s : op == ">" ? a.s > b.s : op == "<=" ? a.s <= b.s : a.s >= b.s;
            else {
                const double x = number(a), y = number(b);
                r = op == "<" ? x < y : op == ">" ? x > y : op == "<=" ? x <= y : x >= y;
            }
            return Value::boolean(r);
        }
        const double x = number(a), y = number(b);
        if (op == "+") return Value::num(x + y);
        if (op == "-") return Value::num(x - y);
        if (op == "*") return Value::num(x * y);
        if (op == "/") return Value::num(x / y);
        if (op == "//") return Value::num(std::floor(x / y));
        if (op == "%") return Value::num(x - std::floor(x / y) * y);
        if (op == "^") return Value::num(std::pow(x, y));
        fail("unsupported operator " + op);
    }

    Value table()
    {
        expect("{");
        auto t = std::make_shared<Table>();
        double pos = 1;
        while (!is_op("}")) {
            if (is_op("[")) {
                lx_.take();
                Value k = expr();
                expect("]");
                expect("=");
                t->set(k, expr());
            } else if (lx_.peek().t == TName && !is_kw("function") && !is_kw("nil") && !is_kw("true") &&
                       !is_kw("false") && !is_kw("not") && peek_assign()) {
                const std::string k = name();
                expect("=");
                t->set(Value::str(k), expr());
            } else {
                t->set(Value::num(pos++), expr());
            }
            if (is_op(",") || is_op(";")) lx_.take();
            else if (!is_op("}")) fail("'}' expected");
        }
        lx_.take();
        Value v;
        v.k = Value::Tab;
        v.t = t;
        return v;
    }
    // is the name at the cursor followed by a single '=' (a field key)?
    bool peek_assign()
    {
        Lexer save = lx_;
        lx_.take();
        const bool r = is_op("=");
        lx_ = save;
        return r;
    }

    // a function value: `(` params `)` then the body, skipped to its `end`
    Value function_value()
    {
        expect("(");
        std::vector<std::string> params;
        while (!is_op(")")) {
            if (is_op("...")) lx_.take();
            else params.push_back(name());
            if (is_op(",")) lx_.take();
            else if (!is_op(")")) fail("')' expected");
        }
        lx_.take();
        auto fd = std::make_shared<FuncDef>(FuncDef{std::move(params), lx_, locals_});
        int depth = 1;
        while (depth > 0) {
            const Token t = lx_.take();
            if (t.t == TEOF) fail("'end' expected");
            if (opens(t)) ++depth;
            else if (closes(t)) --depth;
        }
        Value v;
        v.k = Value::Fun;
        v.f = fd;
        return v;
    }

    Value call(const Value& f, const std::vector<Value>& args) const
    {
        if (f.k != Value::Builtin) fail("only the math library can be called");
        auto a = [&](size_t i) { return i < args.size() ? number(args[i]) : 0.0; };
        const std::string& n = f.s;
        if (n == "sqrt") return Value::num(std::sqrt(a(0)));
        if (n == "sin") return Value::num(std::sin(a(0)));
        if (n == "cos") return Value::num(std::cos(a(0)));
        if (n == "tan") return Value::num(std::tan(a(0)));
        if (n == "asin") return Value::num(std::asin(a(0)));
        if (n == "acos") return Value::num(std::acos(a(0)));
        if (n == "atan") return Value::num(args.size() > 1 ? std::atan2(a(0), a(1)) : std::atan(a(0)));
        if (n == "exp") return Value::num(std::exp(a(0)));
        if (n == "log") return Value::num(args.size() > 1 ? std::log(a(0)) / std::log(a(1)) : std::log(a(0)));
        if (n == "abs") return Value::num(std::fabs(a(0)));
        if (n == "floor") return Value::num(std::floor(a(0)));
        if (n == "ceil") return Value::num(std::ceil(a(0)));
        if (n == "fmod") return Value::num(std::fmod(a(0), a(1)));
        if (n == "min" || n == "max") {
            if (args.empty()) fail("math." + n + " needs arguments");
            double r = a(0);
            for (size_t i = 1; i < args.size(); ++i) r = n == "min" ? std::min(r, a(i)) : std::max(r, a(i));
            return Value::num(r);
        }
        fail("unknown function math." + n);
    }

    Value primary()
    {
        const Token& t = lx_.peek();
        if (t.t == TNum) return Value::num(lx_.take().n);
        if (t.t == TStr) return Value::str(lx_.take().s);
        if (is_op("{")) return table();
        if (is_op("(")) {
            lx_.take();
            Value v = expr();
            expect(")");
            return suffixes(v, "(...)");
        }
        if (t.t == TName) {
            const std::string n = lx_.take().s;
            if (n == "nil") return Value{};
            if (n == "true") return Value::boolean(true);
            if (n == "false") return Value::boolean(false);
            if (n == "function") return function_value();
            return suffixes(lookup(n), n);
        }
        fail("unexpected symbol '" + t.s + "'");
    }

    Value suffixes(Value v, std::string path)
    {
        for (;;) {
            if (is_op(".")) {
                lx_.take();
                const std::string k = name();
                if (v.k != Value::Tab) fail("indexing a non-table value ('" + path + "." + k + "')");
                v = v.t->get(k);
                path += "." + k;
            } else if (is_op("[")) {
                lx_.take();
                Value k = expr();
                expect("]");
                if (v.k != Value::Tab) fail("indexing a non-table value ('" + path + "[...]')");
                v = v.t->get(k);
                path += "[...]";
            } else if (is_op("(")) {
                lx_.take();
                std::vector<Value> args;
                if (!is_op(")")) args = explist();
                expect(")");
                if (v.k == Value::Builtin) {
                    v = call(v, args);
                } else if (v.k == Value::Fun) {
                    std::vector<Value> r = invoke(v, args);
                    v = r.empty() ? Value{} : r[0];
                } else {
                    fail("cannot call " + path + ": not a function");
                }
                path += "(...)";
            } else {
                return v;
            }
        }
    }
};

// ------------------------------------------------------------ config view
const Value* field(const Value& t, const char* k, Value& tmp)
{
    if (t.k != Value::Tab) return nullptr;
    tmp = t.t->get(std::string(k));
    return tmp.k == Value::Nil ? nullptr : &tmp;
}

double num_or(const Value& t, const char* k, double d)
{
    Value tmp;
    const Value* v = field(t, k, tmp);
    return v && v->k == Value::Num ? v->n : d;
}

std::string str_or(const Value& t, const char* k, const std::string& d)
{
    Value tmp;
    const Value* v = field(t, k, tmp);
    return v && v->k == Value::Str ? v->s : d;
}

Value sub(const Value& t, const char* k)
{
    Value tmp;
    const Value* v = field(t, k, tmp);
    return v ? *v : Value{};
}

double idx_or(const Value& t, int i, double d)
{
    if (t.k != Value::Tab) return d;
    const Value v = t.t->get(i);
    return v.k == Value::Num ? v.n : d;
}

// manufactured_solution.type = "lua" (mms.cpp:65-84, lua_mms.hpp:12-79): the
// functions call, ddt, grad, lap of (time, loc), loc = {x, y, z}, called
// through shx_mms_fns.  An error in a function gives NaN (the statistics then
// end the run) and keeps the first message.
struct LuaMms {
    Eval* ev = nullptr;
    Value call, ddt, grad, lap;
    std::string err;

    std::vector<Value> run(const Value& f, double t, const double* loc)
    {
        auto tab = std::make_shared<Table>();
        for (int a = 0; a < 3; ++a) tab->set(Value::num(a + 1), Value::num(loc[a]));
        Value lv;
        lv.k = Value::Tab;
        lv.t = tab;
        return ev->invoke(f, {Value::num(t), lv});
    }
    double scalar(const Value& f, double t, const double* loc)
    {
        try {
            const std::vector<Value> r = run(f, t, loc);
            if (r.empty() || r[0].k != Value::Num) throw std::runtime_error("a manufactured_solution function did not return a number");
            return r[0].n;
        } catch (const std::exception& e) {
            if (err.empty()) err = e.what();
            return std::numeric_limits<double>::quiet_NaN();
        }
    }
    static double cb_call(void* c, double t, const double* loc)
    {
        LuaMms* m = static_cast<LuaMms*>(c);
        return m->scalar(m->call, t, loc);
    }
    static double cb_ddt(void* c, double t, const double* loc)
    {
        LuaMms* m = static_cast<LuaMms*>(c);
        return m->scalar(m->ddt, t, loc);
    }
    static double cb_lap(void* c, double t, const double* loc)
    {
        LuaMms* m = static_cast<LuaMms*>(c);
        return m->scalar(m->lap, t, loc);
    }
    static void cb_grad(void* c, double t, const double* loc, double* g)
    {
        LuaMms* m = static_cast<LuaMms*>(c);
        try {
            const std::vector<Value> r = m->run(m->grad, t, loc);
            for (int a = 0; a < 3; ++a) {
                if (a >= (int)r.size() || r[(size_t)a].k != Value::Num)
                    throw std::runtime_error("manufactured_solution.grad must return three numbers");
                g[a] = r[(size_t)a].n;
            }
        } catch (const std::exception& e) {
            if (m->err.empty()) m->err = e.what();
            g[0] = g[1] = g[2] = std::numeric_limits<double>::quiet_NaN();
        }
    }
};

}  // namespace

struct shoccs_config {
    // the chunk and its evaluator live as long as the configuration: the
    // manufactured solution's functions run in them
    std::string src;
    std::unique_ptr<Eval> ev;
    LuaMms mms;
    shx_mms_fns mms_fns{};
    shoccs_problem_desc desc{};
    std::vector<double> spheres;
    std::vector<double> shapes8;  // shoccs_problem_desc.shapes8
    int eigenvalues = 0;          // system.type = "eigenvalues"
    std::vector<int32_t> shape_bc;
    std::vector<double> params;
    shoccs_run_desc run{};
    int fused = 0;
    // field_io::from_lua (src/io/field_io.cpp:67-106)
    shoccs_io_desc io{};
    std::string io_dir = "io", io_xmf = "view.xmf";
    bool has_io = false;
};

namespace {

int32_t bc_of(const std::string& s, int extent, std::string& err, const char* where)
{
    // boundaries.cpp:17-39: a Dirichlet or Neumann side of an axis of extent 1
    // is taken as floating
    if (s == "dirichlet") return extent == 1 ? SHX_BC_FLOATING : SHX_BC_DIRICHLET;
    if (s == "floating") return SHX_BC_FLOATING;
    if (s == "neumann") return extent == 1 ? SHX_BC_FLOATING : SHX_BC_NEUMANN;
    // extension for bodies of the constant-velocity wave (resolved per cut point)
    if (s == "inflow" && std::string(where) == "shapes.boundary_condition") return SHX_BC_INFLOW;
    err = std::string("incorrect boundary condition '") + s + "' at " + where;
    return -1;
}

int build(const Value& sim, shoccs_config& c, std::string& err)
{
    if (sim.k != Value::Tab) {
        err = "a global table `simulation` is required";
        return SHX_EINVAL;
    }
    shoccs_problem_desc& d = c.desc;
    // mesh (cartesian.cpp:45-89)
    const Value mesh = sub(sim, "mesh");
    const Value ie = sub(mesh, "index_extents"), db = sub(mesh, "domain_bounds");
    if (ie.k != Value::Tab || db.k != Value::Tab) {
        err = "simulation.mesh.index_extents and simulation.mesh.domain_bounds are required";
        return SHX_EINVAL;
    }
    for (int a = 0; a < 3; ++a) d.n[a] = (int32_t)idx_or(ie, a + 1, 1.0);
    const Value lo = sub(db, "min"), hi = sub(db, "max");
    for (int a = 0; a < 3; ++a) d.min[a] = idx_or(lo, a + 1, 0.0);
    for (int a = 0; a < 3; ++a) d.max[a] = hi.k == Value::Tab ? idx_or(hi, a + 1, d.min[a]) : idx_or(db, a + 1, d.min[a]);
    if (hi.k != Value::Tab && db.t->get(1).k == Value::Nil) {
        err = "domain_bounds.max = {...} or domain_bounds = {...} must be specified";
        return SHX_EINVAL;
    }
    // domain_boundaries (boundaries.cpp:41-76): floating by default
    const Value dbc = sub(sim, "domain_boundaries");
    const char* sides[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
    for (int f = 0; f < 6; ++f) {
        d.grid_bc[f] = SHX_BC_FLOATING;
        const std::string s = str_or(dbc, sides[f], "");
        if (!s.empty()) {
            const int32_t b = bc_of(s, d.n[f / 2], err, sides[f]);
            if (b < 0) return SHX_EINVAL;
            d.grid_bc[f] = b;
        }
    }
    // shapes (object_geometry.cpp:244-303, boundaries.cpp:79-87)
    const Value shapes = sub(sim, "shapes");
    if (shapes.k == Value::Tab) {
        for (int i = 1; shapes.t->get(i).k != Value::Nil; ++i) {
            const Value s = shapes.t->get(i);
            const std::string type = str_or(s, "type", "");
            if (type == "sphere") {
                const Value ctr = sub(s, "center");
                double q[8] = {SHOCCS_SPHERE, 0, 0, 0, 0, 0, 0, 0};
                for (int a = 0; a < 3; ++a) q[1 + a] = idx_or(ctr, a + 1, 0.0);
                q[4] = num_or(s, "radius", 0.0);
                c.spheres.insert(c.spheres.end(), q + 1, q + 5);
                c.shapes8.insert(c.shapes8.end(), q, q + 8);
            } else if (type == "yz_rect") {
                // object_geometry.cpp:268-290: corners default to the x = min
                // plane, one cell beyond the domain in y and z; psi places the
                // plane (1 - psi) h inside the left (normal > 0) or right wall
                const double h = (d.max[0] - d.min[0]) / (d.n[0] - 1);
                const Value lcv = sub(s, "lower_corner"), ucv = sub(s, "upper_corner");
                double lc[3] = {idx_or(lcv, 1, d.min[0]), idx_or(lcv, 2, d.min[1] - h), idx_or(lcv, 3, d.min[2] - h)};
                double uc[3] = {idx_or(ucv, 1, d.min[0]), idx_or(ucv, 2, d.max[1] + h), idx_or(ucv, 3, d.max[2] + h)};
                const double nrm = num_or(s, "normal", 1.0);
                Value tmp;
                if (const Value* psi = field(s, "psi", tmp)) {
                    if (psi->k != Value::Num) {
                        err = "shapes.psi must be a number";
                        return SHX_EINVAL;
                    }
                    lc[0] = uc[0] = nrm > 0.0 ? d.min[0] + (1 - psi->n) * h : d.max[0] - (1 - psi->n) * h;
                }
                // make_yz_rect (rect.cpp:8-31): plane at lc[0], bounds in (y, z)
                const double q[8] = {SHOCCS_RECT_X, lc[0], lc[1], lc[2], uc[1], uc[2], nrm > 0 ? 1.0 : -1.0, 0};
                c.spheres.insert(c.spheres.end(), 4, 0.0);
                c.shapes8.insert(c.shapes8.end(), q, q + 8);
            } else {
                err = "shape type must be one of: sphere, yz_rect (got '" + type + "')";
                return SHX_EINVAL;
            }
            const std::string b = str_or(s, "boundary_condition", "floating");
            const int32_t bc = bc_of(b, 2, err, "shapes.boundary_condition");
            if (bc < 0) return SHX_EINVAL;
            c.shape_bc.push_back(bc);
        }
    }
    d.nshapes = (int32_t)c.shape_bc.size();
    // scheme (stencil.cpp:8-106)
    const Value sch = sub(sim, "scheme");
    if (sch.k != Value::Tab) {
        err = "simulation.scheme must be specified";
        return SHX_EINVAL;
    }
    const int order = (int)num_or(sch, "order", 1.0);
    const std::string stype = str_or(sch, "type", "");
    const Value alpha = sub(sch, "alpha");
    d.alpha[0] = idx_or(alpha, 1, 0.0);
    d.alpha[1] = idx_or(alpha, 2, 0.0);
    if (order == 2 && stype == "E2") d.scheme = SHOCCS_E2_2;
    else if (order == 2 && stype == "E4") d.scheme = SHOCCS_E4_2;
    else if (order == 1 && stype == "E4") d.scheme = SHOCCS_E4_1;
    else if (order == 1 && stype == "E4u") d.scheme = SHOCCS_E4U_1;
    else if (order == 1 && stype == "E2-poly") {
        // stencil.cpp:79-99: one alpha list (floating 6, Dirichlet 3,
        // interpolant 4, the list split where it would be), else the three
        // named lists; missing entries are zero (copy_zero_padded)
        d.scheme = SHOCCS_E2P_1;
        for (int i = 0; i < 13; ++i) d.alpha[i] = 0.0;
        if (alpha.k == Value::Tab && alpha.t->get(1).k != Value::Nil) {
            for (int i = 0; i < 13; ++i) d.alpha[i] = idx_or(alpha, i + 1, 0.0);
        } else {
            const Value fa = sub(sch, "floating_alpha"), da = sub(sch, "dirichlet_alpha"),
                        ia = sub(sch, "interpolant_alpha");
            for (int i = 0; i < 6; ++i) d.alpha[i] = idx_or(fa, i + 1, 0.0);
            for (int i = 0; i < 3; ++i) d.alpha[6 + i] = idx_or(da, i + 1, 0.0);
            for (int i = 0; i < 4; ++i) d.alpha[9 + i] = idx_or(ia, i + 1, 0.0);
        }
    } else {
        err = "scheme.order/type = " + std::to_string(order) + " / " + stype +
              " not built (E2 and E4 second derivatives, E4, E4u and E2-poly first derivatives)";
        return SHX_EUNSUPPORTED;
    }
    // system (system.cpp:103-132)
    const Value sys = sub(sim, "system");
    const std::string type = str_or(sys, "type", "");
    if (type == "heat") {
        // heat.cpp:82 and the manufactured solution (mms.cpp:33-95)
        d.kind = SHX_SYS_HEAT;
        c.params.push_back(num_or(sys, "diffusivity", 1.0));
        const Value ms = sub(sim, "manufactured_solution");
        std::string mt = str_or(ms, "type", "");
        for (char& ch : mt) ch = (char)std::tolower((unsigned char)ch);
        if (mt == "gaussian" || ms.k != Value::Tab) {
            c.params.push_back(SHX_MMS_GAUSS);
            std::vector<double> g;
            int ng = 0;
            if (ms.k == Value::Tab)
                for (int i = 1; ms.t->get(i).k != Value::Nil; ++i, ++ng) {
                    const Value t = ms.t->get(i);
                    const Value cc = sub(t, "center"), vv = sub(t, "variance");
                    for (int a = 0; a < 3; ++a) g.push_back(idx_or(cc, a + 1, 0.0));
                    for (int a = 0; a < 3; ++a) g.push_back(idx_or(vv, a + 1, 0.0));
                    g.push_back(num_or(t, "amplitude", 0.0));
                    g.push_back(num_or(t, "frequency", 0.0));
                }
            c.params.push_back(ng);
            c.params.insert(c.params.end(), g.begin(), g.end());
        } else if (mt == "sin_product") {
            // this library's extension: sin(2 pi x) sin(2 pi y) [sin(2 pi z)] exp(-t)
            c.params.push_back(SHX_MMS_SINPROD);
        } else if (mt == "lua") {
            // mms.cpp:65-84: call, ddt, grad, div and lap are all required
            const char* fns[] = {"call", "ddt", "grad", "div", "lap"};
            for (const char* f : fns)
                if (sub(ms, f).k != Value::Fun) {
                    err = "`lua` manufactured solution requires call, ddt, grad, div, lap functions";
                    return SHX_EINVAL;
                }
            c.mms.ev = c.ev.get();
            c.mms.call = sub(ms, "call");
            c.mms.ddt = sub(ms, "ddt");
            c.mms.grad = sub(ms, "grad");
            c.mms.lap = sub(ms, "lap");
            c.mms_fns = shx_mms_fns{&c.mms, &LuaMms::cb_call, &LuaMms::cb_ddt, &LuaMms::cb_grad, &LuaMms::cb_lap};
            c.desc.mms_fns = &c.mms_fns;
            c.params.push_back(SHX_MMS_HOST);
        } else {
            err = "manufactured_solution.type = '" + mt + "': expected gaussian, lua or sin_product";
            return SHX_EINVAL;
        }
    } else if (type == "scalar wave") {
        const Value vel = sub(sys, "velocity");
        const double max_error = num_or(sys, "max_error", 100.0);  // scalar_wave.cpp:137
        if (vel.k == Value::Tab) {
            // this library's extension: constant velocity, translated gaussian
            d.kind = SHX_SYS_WAVE_LINEAR;
            const Value ic = sub(sys, "initial_condition");
            const Value x0 = sub(ic, "center");
            for (int a = 0; a < 3; ++a) c.params.push_back(idx_or(vel, a + 1, 0.0));
            for (int a = 0; a < 3; ++a) c.params.push_back(idx_or(x0, a + 1, 0.0));
            c.params.push_back(num_or(ic, "width", 0.05));
            c.params.push_back(num_or(ic, "amplitude", 1.0));
            c.params.push_back(max_error);
        } else {
            // scalar_wave.cpp:137-172: system.center/radius, else the first sphere
            d.kind = SHX_SYS_WAVE_RADIAL;
            double ctr[3] = {0, 0, 0}, rad = 0;
            const Value sc = sub(sys, "center");
            if (sc.k == Value::Tab && sub(sys, "radius").k == Value::Num) {
                for (int a = 0; a < 3; ++a) ctr[a] = idx_or(sc, a + 1, 0.0);
                rad = num_or(sys, "radius", 0.0);
            } else if (!c.shapes8.empty() && c.shapes8[0] == SHOCCS_SPHERE) {
                for (int a = 0; a < 3; ++a) ctr[a] = c.spheres[a];
                rad = c.spheres[3];
            } else {
                err = "a system.center / system.radius must be specified for scalar_wave";
                return SHX_EINVAL;
            }
            c.params.assign(ctr, ctr + 3);
            c.params.push_back(rad);
            c.params.push_back(max_error);
        }
    } else if (type == "eigenvalues") {
        // hyperbolic_eigenvalues (hyperbolic_eigenvalues.cpp:39-52): the
        // mesh, boundaries and scheme only; its statistic is computed from the
        // x derivative (shoccs_amd.spectral).  The operators are carried by a
        // constant-velocity wave along x, which is never stepped.
        d.kind = SHX_SYS_WAVE_LINEAR;
        c.eigenvalues = 1;
        const double p[9] = {1.0, 0.0, 0.0, d.min[0], d.min[1], d.min[2], 1.0, 1.0, 100.0};
        c.params.assign(p, p + 9);
    } else {
        err = "system.type '" + type + "' not built (heat, scalar wave, eigenvalues)";
        return SHX_EUNSUPPORTED;
    }
    d.spheres = c.spheres.data();
    d.shapes8 = c.shapes8.empty() ? nullptr : c.shapes8.data();
    d.shape_bc = c.shape_bc.data();
    d.params = c.params.data();
    d.nparams = (int32_t)c.params.size();
    // integrator (integrator.cpp:25-48); `form = "fused"` is this library's
    // fused RK4
    const Value integ = sub(sim, "integrator");
    const std::string it = str_or(integ, "type", "rk4");
    c.fused = str_or(integ, "form", "") == "fused";
    if (it == "rk4") c.run.integrator = c.fused ? SHOCCS_RK4_FUSED : SHOCCS_RK4;
    else if (it == "euler") c.run.integrator = SHOCCS_EULER;
    else {
        err = "integrator.type must be one of: [rk4, euler]";
        return SHX_EINVAL;
    }
    // step_controller (step_controller.cpp:7-28)
    const Value sc = sub(sim, "step_controller");
    const Value cfl = sub(sc, "cfl");
    const double ms = num_or(sc, "max_step", (double)std::numeric_limits<int32_t>::max());
    c.run.max_step = ms >= (double)std::numeric_limits<int32_t>::max() ? std::numeric_limits<int32_t>::max()
                                                                         : (int32_t)ms;
    c.run.max_time = num_or(sc, "max_time", std::numeric_limits<double>::max());
    c.run.min_dt = num_or(sc, "min_dt", 1e-6);
    c.run.h_cfl = num_or(cfl, "hyperbolic", 1.0);
    c.run.p_cfl = num_or(cfl, "parabolic", 1.0);
    // io (field_io.cpp:67-106): write_every_step or write_every_time, dir,
    // suffix_length, xdmf_filename
    const Value io = sub(sim, "io");
    c.io_dir = str_or(io, "dir", "io");
    c.io_xmf = str_or(io, "xdmf_filename", "view.xmf");
    c.io.suffix_length = (int32_t)num_or(io, "suffix_length", 6.0);
    c.io.write_every_step = (int32_t)num_or(io, "write_every_step", 0.0);
    c.io.write_every_time = num_or(io, "write_every_time", 0.0);
    c.has_io = c.io.write_every_step > 0 || c.io.write_every_time > 0.0;
    return SHX_OK;
}

void put_err(const std::string& m, char* err, int64_t errlen)
{
    if (!err || errlen <= 0) return;
    std::strncpy(err, m.c_str(), (size_t)errlen - 1);
    err[errlen - 1] = 0;
}

}  // namespace

extern "C" {

int shoccs_config_parse(const char* text, shoccs_config** out, char* err, int64_t errlen)
{
    if (!text || !out) return SHX_EINVAL;
    *out = nullptr;
    try {
        auto c = std::make_unique<shoccs_config>();
        c->src = text;
        c->ev = std::make_unique<Eval>(c->src);
        c->ev->run();
        std::string msg;
        const int rc = build(c->ev->global("simulation"), *c, msg);
        if (rc != SHX_OK) {
            put_err(msg, err, errlen);
            return rc;
        }
        *out = c.release();
        return SHX_OK;
    } catch (const std::exception& e) {
        put_err(e.what(), err, errlen);
        return SHX_EINVAL;
    }
}

int shoccs_config_desc(const shoccs_config* c, shoccs_problem_desc* desc, shoccs_run_desc* run)
{
    if (!c) return SHX_EINVAL;
    if (desc) *desc = c->desc;
    if (run) *run = c->run;
    return SHX_OK;
}

int shoccs_config_io(const shoccs_config* c, shoccs_io_desc* io)
{
    if (!c || !io) return SHX_EINVAL;
    *io = c->io;
    io->dir = c->io_dir.c_str();
    io->xdmf_filename = c->io_xmf.c_str();
    return c->has_io ? 1 : 0;
}

int shoccs_config_eigenvalues(const shoccs_config* c)
{
    return c ? c->eigenvalues : SHX_EINVAL;
}

const char* shoccs_config_error(const shoccs_config* c)
{
    return c ? c->mms.err.c_str() : "";
}

int shoccs_config_destroy(shoccs_config* c)
{
    if (!c) return SHX_EINVAL;
    delete c;
    return SHX_OK;
}

}  // extern "C"
