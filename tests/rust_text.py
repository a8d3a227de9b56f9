"""Evaluates small integer functions of the reference's Rust source, read as
text, in Python.

Test infrastructure: it lets a test compare an oracle function with the
reference's own expressions instead of a retyped copy of them. It translates
the subset of Rust that the reference's per-line filter and decision
functions use (crates/rustmpeg-video/src/vvc/filter.rs):

- `let` bindings (with `mut`, tuple patterns, type annotations), closures
  `|a: T, b: T| expr` and `|x: T| -> R { match .. }`;
- assignments to variables and to `line[...]`;
- `if COND { .. }` / `if COND { .. } else { .. }` as statements and as
  expressions, `return;`, `for PAT in ITER { .. }`;
- `match SCRUT { PAT | PAT => EXPR, _ => { ..; EXPR } }` as an expression, and
  as a statement with statement arms;
- the methods `.clamp(a, b)`, `.abs()`, `.min(x)`, `.max(x)`, `.iter()`,
  `.zip(y)`, `.enumerate()`;
- `&&`, `||`, `!`, references (`&m`, `&[..]`, `*line`), `i32::from(x)`,
  `x as T`, integer suffixes.

Rust's precedence for the operators used (`*`, `/`, `+`, `-`, `>>`, `<<`,
comparisons, `&&`, `||`) is Python's, and `>>` on i32 is an arithmetic shift
like Python's, so the translated expressions compute what the Rust ones do
for values that do not overflow i32. Integer `/` becomes Python's floor
division, equal to Rust's truncating division on the non-negative operands
these functions divide.
"""
import re
import types

_METHODS = ("clamp", "abs", "min", "max", "iter", "zip", "enumerate", "map", "collect", "to_vec", "clone", "len",
            "sum", "saturating_sub", "trailing_zeros", "fill", "step_by", "wrapping_sub", "get", "copied",
            "unwrap_or", "div_ceil", "copy_from_slice", "expect", "is_empty", "first", "and_then", "as_ref", "rev",
            "take", "ok_or", "map_err", "signum", "contains", "leading_zeros")


def fn_source(text, name, after=None):
    """(parameter list, body) of `fn name(...) ... { body }` in text, the first
    one past the text `after` when given (a method of one impl block)."""
    if after is not None:
        text = text[text.index(after):]
    m = re.search(r"\bfn\s+%s\s*(<[^>]*>)?\s*\(" % re.escape(name), text)
    if not m:
        raise KeyError(name)
    i = m.end() - 1
    j = _match(text, i)
    params = text[i + 1:j]
    k = text.index("{", j)
    end = _match(text, k)
    return params, text[k + 1:end]


def _match(s, i):
    """Index of the bracket closing the one at s[i]."""
    pairs = {"(": ")", "[": "]", "{": "}"}
    stack = [pairs[s[i]]]
    j = i + 1
    while stack:
        c = s[j]
        if c in pairs:
            stack.append(pairs[c])
        elif c in ")]}":
            assert c == stack[-1], (s[i:j + 1], c)
            stack.pop()
        j += 1
    return j - 1


def _split_top(s, sep):
    """Splits s at `sep` characters outside brackets."""
    out, depth, cur = [], 0, []
    for c in s:
        if c in "([{":
            depth += 1
        elif c in ")]}":
            depth -= 1
        if c == sep and depth == 0:
            out.append("".join(cur))
            cur = []
        else:
            cur.append(c)
    out.append("".join(cur))
    return out


def _receiver_start(s, dot):
    """Start index of the receiver expression that ends just before s[dot]
    (a method call may sit on the next line: `(..)\n    .clamp(..)`)."""
    j = dot
    while j > 0 and s[j - 1].isspace():
        j -= 1
    while True:
        if j > 0 and s[j - 1] in ")]":
            depth, k = 0, j - 1
            while True:
                if s[k] in ")]":
                    depth += 1
                elif s[k] in "([":
                    depth -= 1
                    if depth == 0:
                        break
                k -= 1
            j = k
            continue
        if j > 0 and (s[j - 1].isalnum() or s[j - 1] == "_" or (s[j - 1] == "." and s[j - 2:j] != "..")):
            j -= 1
            continue
        return j


def _methods(s):
    while True:
        m = re.search(r"\.(%s)\(" % "|".join(_METHODS), s)
        if not m:
            return s
        name = m.group(1)
        start = _receiver_start(s, m.start())
        recv = s[start:m.start()]
        open_ = m.end() - 1
        close = _match(s, open_)
        args = s[open_ + 1:close].strip()
        if name == "clamp":
            rep = "_clamp(%s, %s)" % (recv, args)
        elif name == "abs":
            rep = "_abs(%s)" % recv
        elif name in ("min", "max"):
            rep = "_%s(%s, %s)" % (name, recv, args)
        elif name == "iter":
            rep = recv
        elif name == "zip":
            rep = "_zip(%s, %s)" % (recv, args)
        elif name == "enumerate":
            rep = "_enumerate(%s)" % recv
        elif name == "map":
            cm = re.match(r"\|([^|]*)\|(.*)$", args, re.S)
            var = cm.group(1).split(":")[0].strip()
            rep = "[(%s) for %s in (%s)]" % (expr(cm.group(2)), var, recv)
        elif name in ("collect",):
            rep = recv
        elif name in ("to_vec", "clone"):
            rep = "_list(%s)" % recv
        elif name == "len":
            rep = "_len(%s)" % recv
        elif name == "fill":
            rep = "_fill(%s, %s)" % (recv, args)
        elif name == "step_by":
            rep = "_step_by(%s, %s)" % (recv, args)
        elif name == "wrapping_sub":
            rep = "(%s - (%s))" % (recv, args)  # used where the difference does not wrap
        elif name == "get":
            rep = "_get(%s, %s)" % (recv, args)  # Option: None past the end
        elif name == "copied":
            rep = recv
        elif name == "unwrap_or":
            rep = "_unwrap_or(%s, %s)" % (recv, args)
        elif name == "copy_from_slice":
            rep = "_copy_into(%s, %s)" % (recv, args)
        elif name == "contains":
            rep = "((%s) in (%s))" % (args.lstrip("&"), recv)
        elif name == "rev":
            rep = "_rev(%s)" % recv
        elif name == "take":
            rep = "_take(%s, %s)" % (recv, args)
        elif name == "ok_or":
            rep = "_unwrap_opt(%s)" % recv
        elif name == "map_err":
            rep = recv
        elif name == "signum":
            rep = "_signum(%s)" % recv
        elif name == "first":
            rep = "_first(%s)" % recv
        elif name == "and_then":
            rep = "_and_then(%s, %s)" % (recv, expr(args))
        elif name == "as_ref":
            rep = recv
        elif name == "is_empty":
            rep = "(_len(%s) == 0)" % recv
        elif name == "expect":
            rep = recv
        elif name == "div_ceil":
            rep = "(-(-(%s) // (%s)))" % (recv, args)
        elif name == "sum":
            rep = "_sum(%s)" % recv
        elif name == "saturating_sub":
            rep = "_max(%s - (%s), 0)" % (recv, args)
        elif name == "leading_zeros":
            rep = "(32 - (%s).bit_length())" % recv  # of a non-negative u32
        else:
            rep = "((%s) & -(%s)).bit_length() - 1" % (recv, recv)
        s = s[:start] + rep + s[close + 1:]


def _block_value(body):
    """Python expression for a Rust block `{ stmts; expr }` used as a value
    (assertions dropped; the block must end in its value)."""
    stmts = [t.strip() for t in _split_top(body, ";")]
    stmts = [t for t in stmts if t and not t.startswith("debug_assert")]
    lets = []
    for t in stmts[:-1]:
        # `let name = e;` before the value: an assignment expression
        m = re.match(r"^let\s+(?:mut\s+)?(\w+)\s*(?::[^=]+)?=(.+)$", t, re.S)
        assert m, stmts
        lets.append("(%s := (%s))" % (m.group(1), expr(m.group(2))))
    if not lets:
        return expr(stmts[-1])
    return "(%s, (%s))[-1]" % (", ".join(lets), expr(stmts[-1]))


def _match_expr(s, at):
    """(end, python) of the `match` expression starting at s[at]."""
    brace = s.index("{", at)
    scrut = s[at + 5:brace].strip()
    close = _match(s, brace)
    arms, default = [], None
    rest = s[brace + 1:close].strip()
    while rest:
        arrow = rest.index("=>")
        pat = rest[:arrow].strip()
        rest = rest[arrow + 2:].lstrip()
        if rest.startswith("{"):
            e = _match(rest, 0)
            val = _block_value(rest[1:e])
            rest = rest[e + 1:].lstrip()
            if rest.startswith(","):
                rest = rest[1:].lstrip()
        else:
            parts = _split_top(rest, ",")
            val = expr(parts[0])
            rest = ",".join(parts[1:]).lstrip()
        if pat == "_":
            default = val
        else:
            arms.append(([p.strip() for p in _split_top(pat, "|")], val))
    sv = expr(scrut)
    out = default if default is not None else "None"
    for alts, val in reversed(arms):
        cond = " or ".join("(%s) == (%s)" % (sv, expr(a)) for a in alts)
        out = "((%s) if (%s) else %s)" % (val, cond, out)
    return close + 1, out


def _if_expr(s, at):
    """(end, python) of the `if .. { } else { }` expression starting at s[at]."""
    brace = _top_brace(s, at + 2)
    cond = s[at + 2:brace]
    close = _match(s, brace)
    then = _block_value(s[brace + 1:close])
    j = close + 1
    while s[j].isspace():
        j += 1
    assert s.startswith("else", j), "if without else used as a value"
    j += 4
    while s[j].isspace():
        j += 1
    if s.startswith("if", j):
        end, other = _if_expr(s, j)
    else:
        end = _match(s, j) + 1
        other = "(%s)" % _block_value(s[j + 1:end - 1])
    return end, "((%s) if (%s) else %s)" % (then, expr(cond), other)


def _top_brace(s, start):
    """Index of the first `{` at bracket depth 0 from start (ends an `if`
    condition or a `for` header)."""
    depth = 0
    for j in range(start, len(s)):
        c = s[j]
        if c == "{" and depth == 0:
            return j
        if c in "([":
            depth += 1
        elif c in ")]":
            depth -= 1
    raise ValueError(s)


def expr(s):
    """Python source of one Rust expression."""
    s = s.strip()
    stash = []
    # string literals first, so that words inside them are not taken for keywords
    while True:
        m = re.search(r'"(?:[^"\\]|\\.)*"', s)
        if not m:
            break
        s = s[:m.start()] + " __py%d__ " % len(stash) + s[m.end():]
        stash.append(m.group(0))
    # `match` / `if` expressions become Python conditionals, held aside under
    # placeholder names while the rest is translated
    while True:
        # a closure with a block body passed as an argument: |x, y| { stmts; value }
        m = re.search(r"(?<=[(,])\s*\|([^|]*)\|\s*\{", s)
        if not m:
            break
        brace = m.end() - 1
        close = _match(s, brace)
        params = ", ".join(p.split(":")[0].strip().lstrip("&") for p in m.group(1).split(",") if p.strip())
        py = "(lambda %s: %s)" % (params, _block_value(s[brace + 1:close]))
        s = s[:m.start()] + " __py%d__ " % len(stash) + s[close + 1:]
        stash.append(py)
    while True:
        m = re.search(r"\b(match|if)\b", s)
        if not m:
            break
        end, py = (_match_expr if m.group(1) == "match" else _if_expr)(s, m.start())
        s = s[:m.start()] + " __py%d__ " % len(stash) + s[end:]
        stash.append(py)
    s = s.replace("&&", " and ").replace("||", " or ")
    s = re.sub(r"\.sum::<\w+>\(\)", ".sum()", s)
    s = _vec_macro(s)
    s = re.sub(r"\[([^\[\];]+);\s*([^\[\];]+)\]", r"([\1] * (\2))", s)  # [v; n]
    s = _slices(s)
    m = re.match(r"^\s*\|([^|]*)\|", s)
    if m:
        # a closure: |params| body  or  |params| -> T { body }
        params = ", ".join(p.split(":")[0].strip().lstrip("&") for p in m.group(1).split(",") if p.strip())
        rest = s[m.end():].lstrip()
        if rest.startswith("->"):
            brace = rest.index("{")
            body = _block_value(rest[brace + 1:_match(rest, brace)])
        else:
            body = expr(rest)
        s = s[:m.start()] + "(lambda %s: %s)" % (params, body)
    else:
        s = re.sub(r"&mut\s+", "", s)
        s = re.sub(r"\bOk\(", "(", s)  # the functions raise _RustErr instead of returning Err
        s = re.sub(r"\)\?", ")", s)
        s = re.sub(r"\b([A-Z]\w*)::([A-Z]\w*)\b(?!\s*\()", r"\1_\2", s)  # Enum::Variant -> Enum_Variant
        s = re.sub(r",\s*\)", ")", s)  # trailing comma of a call (Rust has no 1-tuples here)
        s = re.sub(r"!(?!=)", " not ", s)
        s = re.sub(r"&(?=[\w\[])", "", s)
        s = re.sub(r"(?<![\w\s)\]])\s*\*(line)\b", r"_list(\1)", s)  # `*line`: a copy of the array
        s = re.sub(r"^\s*\*(\w+)\s*$", r"_deref(\1)", s)  # `*x` as a whole value
        s = s.replace("usize::MAX", "_USIZE_MAX")
        s = re.sub(r"\b(?:i32|i64|u32|usize|u8|i8)::from\(", "(", s)
        s = re.sub(r"\b(?:super::)?[a-z_]+::(?=[A-Z][A-Z0-9_]+\b)", "", s)  # module path of a constant
        s = re.sub(r"\b(?:super::)?(?:filter|inter|predict|dct|transform|lfnst|tr_tables)::(?=[a-z_]\w*\()", "", s)
        s = re.sub(r"\btrue\b", "True", s)
        s = re.sub(r"\bfalse\b", "False", s)
        s = re.sub(r"\s+as\s+(?:usize|isize|i32|i64|u32|u8|i8|u16)\b", "", s)
        s = re.sub(r"\b(\d+)(?:i32|i64|u32|u64|usize|isize|u8|i8|u16|i16)\b", r"\1", s)
        s = re.sub(r"(?<![/*])/(?![/*=])", "//", s)  # integer division (the functions divide non-negative values)
        s = _ranges(s)
        s = _methods(s)
    for k in range(len(stash) - 1, -1, -1):
        s = s.replace("__py%d__" % k, stash[k])
    return s


def _vec_macro(s):
    """vec![v; n] -> [v] * (n)."""
    while "vec![" in s:
        i = s.index("vec![")
        close = _match(s, i + 4)
        v, n = _split_top(s[i + 5:close], ";")
        s = s[:i] + "([%s] * (%s))" % (v.strip(), n.strip()) + s[close + 1:]
    return s


def _slices(s):
    """x[..=e] -> x[:e + 1], x[..e] -> x[:e], x[a..b] -> x[a:b]."""
    out, i = [], 0
    while i < len(s):
        c = s[i]
        if c == "[" and i > 0 and (s[i - 1].isalnum() or s[i - 1] in "_)]"):
            close = _match(s, i)
            inner = s[i + 1:close]
            rm = re.match(r"^(.*?)\.\.(=?)(.*)$", inner, re.S) if ".." in inner and _split_top(inner, ",") == [inner] else None
            if rm:
                a, incl, b = rm.group(1).strip(), rm.group(2), rm.group(3).strip()
                b = "(%s) + 1" % b if incl and b else b
                out.append("[%s:%s]" % (a, b))
            else:
                out.append("[" + _slices(inner) + "]")
            i = close + 1
            continue
        out.append(c)
        i += 1
    return "".join(out)


def _ranges(s):
    """a..b / a..=b outside brackets -> range(a, b) / range(a, b + 1); the
    range must be the whole of a parenthesised group or of the string."""
    def conv(t):
        rm = re.match(r"^(.*?)\.\.(=?)(.*)$", t, re.S)
        a, b = rm.group(1).strip(), rm.group(3).strip()
        return "_range(%s, %s)" % (a, "(%s) + 1" % b if rm.group(2) else b)
    if ".." not in s:
        return s
    if _split_top(s, ".") and re.match(r"^[^()\[\]]*\.\.[^()\[\]]*$", s.strip()):
        return conv(s.strip())
    out, i = [], 0
    while i < len(s):
        if s[i] == "(":
            close = _match(s, i)
            inner = s[i + 1:close]
            if ".." in inner and re.match(r"^[^()\[\]]*\.\.[^()\[\]]*$", inner.strip()):
                out.append("(" + conv(inner) + ")")
            else:
                out.append("(" + _ranges(inner) + ")")
            i = close + 1
            continue
        out.append(s[i])
        i += 1
    return "".join(out)


def _struct_literal(stmt):
    """`Name { a, b: e, }` -> _struct(a=a, b=e), or None."""
    m = re.match(r"^([A-Z]\w*)\s*\{(.*)\}$", stmt.strip(), re.S)
    if not m:
        return None
    fields = [f.strip() for f in _split_top(m.group(2), ",") if f.strip()]
    kv = []
    for f in fields:
        if ":" in f:
            k, v = f.split(":", 1)
            kv.append("%s=(%s)" % (k.strip(), expr(v)))
        else:
            kv.append("%s=%s" % (f, f))
    return "_struct(%s)" % ", ".join(kv)


def _pattern(p):
    p = re.sub(r"\bmut\s+", "", p.strip())
    p = re.sub(r"&(?=[\w(])", "", p)
    depth = 0
    for i, c in enumerate(p):  # cut a type annotation `: T` at bracket depth 0
        if c in "([":
            depth += 1
        elif c in ")]":
            depth -= 1
        elif c == ":" and depth == 0:
            return p[:i].strip()
    return p.strip()


def block(body, indent="    ", tail=None):
    """Python statements (as lines) for the Rust statements of body. The
    block's value (its last expression) is returned, or assigned to `tail`
    when given (a block that is the value of a `let`)."""
    out, s = [], body.strip()
    while s:
        if s.startswith("//"):
            s = s[s.index("\n") + 1:].lstrip() if "\n" in s else ""
            continue
        if re.match(r"if\b", s):
            brace = _top_brace(s, 2)
            cond = s[2:brace]
            close = _match(s, brace)
            lm = re.match(r"^\s*let\s+Some\((\w+)\)\s*=(.+)$", cond, re.S)
            if lm:  # if let Some(x) = e
                out.append(indent + "if ((%s := (%s)) is not None):" % (lm.group(1), expr(lm.group(2))))
            else:
                out.append(indent + "if (%s):" % expr(cond))
            out += block(s[brace + 1:close], indent + "    ", tail) or [indent + "    pass"]
            s = s[close + 1:].lstrip()
            while s.startswith("else"):
                s = s[4:].lstrip()
                if re.match(r"if\b", s):  # else if
                    brace = _top_brace(s, 2)
                    cond = s[2:brace]
                    close = _match(s, brace)
                    out.append(indent + "elif (%s):" % expr(cond))
                else:
                    brace = s.index("{")
                    close = _match(s, brace)
                    out.append(indent + "else:")
                out += block(s[brace + 1:close], indent + "    ", tail) or [indent + "    pass"]
                s = s[close + 1:].lstrip()
            continue
        if re.match(r"while\b", s):
            brace = _top_brace(s, 5)
            cond = s[5:brace]
            close = _match(s, brace)
            out.append(indent + "while (%s):" % expr(cond))
            out += block(s[brace + 1:close], indent + "    ") or [indent + "    pass"]
            s = s[close + 1:].lstrip()
            continue
        if re.match(r"match\b", s):
            # a match statement: arms `PAT | PAT => stmt,` or `PAT => { stmts }`
            brace = s.index("{")
            scrut = expr(s[5:brace])
            close = _match(s, brace)
            rest = s[brace + 1:close].strip()
            first = True
            while rest:
                arrow = rest.index("=>")
                pat = rest[:arrow].strip()
                rest = rest[arrow + 2:].lstrip()
                if rest.startswith("{"):
                    e = _match(rest, 0)
                    body = rest[1:e]
                    rest = rest[e + 1:].lstrip()
                else:
                    parts = _split_top(rest, ",")
                    body = parts[0] + ";"
                    rest = ",".join(parts[1:])
                rest = rest.lstrip()
                if rest.startswith(","):
                    rest = rest[1:].lstrip()
                bind = []
                if pat == "_":
                    out.append(indent + "else:")
                else:
                    em = re.match(r"^\w+::(\w+)\((\w+)\)$", pat)
                    if em:  # Enum::Variant(x): values of such enums are ("Variant", x) tuples
                        cond = "(%s)[0] == %r" % (scrut, em.group(1))
                        bind = [indent + "    %s = (%s)[1]" % (em.group(2), scrut)]
                    else:
                        cond = " or ".join("(%s) == (%s)" % (scrut, expr(a)) for a in _split_top(pat, "|"))
                    out.append(indent + ("if (%s):" if first else "elif (%s):") % cond)
                out += bind + (block(body, indent + "    ", tail) or [indent + "    pass"])
                first = False
            s = s[close + 1:].lstrip()
            continue
        if re.match(r"for\b", s):
            brace = _top_brace(s, 3)
            head = s[3:brace]
            pat, it = head.split(" in ", 1)
            close = _match(s, brace)
            body_src = s[brace + 1:close]
            im = re.match(r"^\s*(\w+)\.iter_mut\(\)\s*$", it)
            tm = re.match(r"^\s*(\w+)\.iter_mut\(\)\.take\((.+)\)\.enumerate\(\)\s*$", it, re.S)
            tm = tm or re.match(r"^\s*(\w+)\.iter_mut\(\)\.enumerate\(\)\.take\((.+)\)\s*$", it, re.S)
            zm = re.match(r"^\s*(\w+)\.iter_mut\(\)\.zip\((\w+)\)\s*$", it)
            pm = re.match(r"^\s*\(\s*(\w+)\s*,\s*&?(\w+)\s*\)\s*$", pat)
            if zm and pm:
                # for (d, &x) in out.iter_mut().zip(xs) { *d = e(x); }  ->  over the indices of out
                seq, var = zm.group(1), pm.group(1)
                k = "_i_%s" % var
                out.append(indent + "for %s in _range(0, _min(_len(%s), _len(%s))):" % (k, seq, zm.group(2)))
                out.append(indent + "    %s = %s[%s]" % (pm.group(2), zm.group(2), k))
                body_src = re.sub(r"\*%s\b" % var, "%s[%s]" % (seq, k), body_src)
            elif tm and pm:
                # for (i, v) in xs.iter_mut().take(n).enumerate() { *v = e; }
                seq, idx, var = tm.group(1), pm.group(1), pm.group(2)
                out.append(indent + "for %s in _range(0, _min(%s, _len(%s))):" % (idx, expr(tm.group(2)), seq))
                body_src = re.sub(r"\*%s\b" % var, "%s[%s]" % (seq, idx), body_src)
            elif im:
                # for v in xs.iter_mut() { *v op= e; }  ->  over the indices of xs
                var, seq = _pattern(pat), im.group(1)
                k = "_i_%s" % var
                out.append(indent + "for %s in _range(0, _len(%s)):" % (k, seq))
                body_src = re.sub(r"\*%s\b" % var, "%s[%s]" % (seq, k), body_src)
            else:
                rm = re.match(r"^(.+?)\.\.(=?)(.+)$", it.strip(), re.S)
                if rm and "[" not in rm.group(1) and "(" not in rm.group(1):
                    b = expr(rm.group(3))
                    rng = "_range(%s, %s)" % (expr(rm.group(1)), "(%s) + 1" % b if rm.group(2) else b)
                else:
                    rng = expr(it)
                out.append(indent + "for %s in (%s):" % (_pattern(pat), rng))
            out += block(body_src, indent + "    ")
            s = s[close + 1:].lstrip()
            continue
        parts = _split_top(s, ";")
        stmt = parts[0].strip()
        s = ";".join(parts[1:]).lstrip()
        if stmt.startswith("const "):
            stmt = "let " + stmt[6:]  # a local constant: a binding
        if not stmt or stmt.startswith("debug_assert"):
            continue
        if not parts[1:]:
            lit = _struct_literal(stmt)
            if tail is not None:
                out.append(indent + "%s = (%s)" % (tail, lit or expr(stmt)))
            else:
                out.append(indent + "return (%s)" % (lit or expr(stmt)))  # tail expression
            continue
        if re.match(r"^let\s+_\s*=", stmt):
            continue
        le = re.match(r"^let\s+Some\((\w+)\)\s*=(.+?)\s+else\s*\{(.*)\}$", stmt, re.S)
        if le:  # let Some(x) = e else { continue / return };
            out.append(indent + "%s = (%s)" % (le.group(1), expr(le.group(2))))
            out.append(indent + "if %s is None:" % le.group(1))
            out += block(le.group(3).strip().rstrip(";") + ";", indent + "    ")
            continue
        cm = re.match(r"^([\w\[\]\s\.\-\+\*]+?)\s*(\+|-|\*|<<|>>|\||&|\^)=\s*(.+)$", stmt, re.S)
        cl = re.match(r"^let\s+(\w+)\s*=\s*\|([^|]*)\|\s*->[^{]*\{", stmt)
        if stmt in ("break", "continue"):
            out.append(indent + stmt)
        elif cl and _match(stmt, stmt.index("{", cl.end() - 1)) == len(stmt.rstrip()) - 1:
            # a closure with a statement body: a nested function
            params = ", ".join(q.split(":")[0].strip() for q in cl.group(2).split(",") if q.strip())
            b0 = stmt.index("{", cl.end() - 1)
            out.append(indent + "def %s(%s):" % (cl.group(1), params))
            inner = stmt[b0 + 1:len(stmt.rstrip()) - 1].strip()
            if re.match(r"match\b", inner) and _match(inner, inner.index("{")) == len(inner) - 1:
                out.append(indent + "    return (%s)" % expr(inner))  # a match as the tail expression
            else:
                out += block(inner, indent + "    ")
        elif stmt == "return":
            out.append(indent + "return None")
        elif re.match(r"^return\s+Err\(", stmt):
            out.append(indent + "raise _RustErr(%s)" % expr(stmt[stmt.index("Err(") + 3:]))
        elif stmt.startswith("return "):
            out.append(indent + "return (%s)" % expr(stmt[7:]))
        elif cm:
            out.append(indent + "%s %s= (%s)" % (cm.group(1).strip(), cm.group(2), expr(cm.group(3))))
        elif stmt.startswith("let "):
            lhs, rhs = stmt[4:].split("=", 1)
            try:
                out.append(indent + "%s = (%s)" % (_pattern(lhs), expr(rhs)))
            except AssertionError:
                if not re.match(r"\s*(match|if)\b", rhs):
                    raise
                # arms with statements: the match / if as statements assigning the binding
                out += block(rhs.strip(), indent, _pattern(lhs))
        elif re.match(r"[\w\[\]\s\-\+\*()]+=[^=]", stmt) and not re.match(r"[\w\s]+[<>!=]=", stmt):
            lhs, rhs = stmt.split("=", 1)
            lhs = re.sub(r"\s+as\s+(?:usize|i32|i64|u32|u8|i8|u16)\b", "", lhs.strip())
            out.append(indent + "%s = (%s)" % (lhs, expr(rhs)))
        else:
            out.append(indent + "(%s)" % expr(stmt))
    return out


def _clamp(v, lo, hi):
    assert lo <= hi, (lo, hi)  # Rust's Ord::clamp panics when min > max
    return lo if v < lo else hi if v > hi else v


_USIZE_MAX = (1 << 64) - 1


def _copy_into(dst, src):
    dst[:] = list(src)


def _deref(v):
    return list(v) if isinstance(v, (list, tuple)) else v


class _RustErr(Exception):
    """An Err(...) of the evaluated Rust function."""


def _unwrap_opt(v):
    if v is None:
        raise _RustErr("None")
    return v


def _rev(xs):
    return list(xs)[::-1]


def _take(xs, n):
    return list(xs)[:n]


def _signum(v):
    return (v > 0) - (v < 0)


def _clone(x):
    return x.clone() if hasattr(x, "clone") else list(x)


def _first(xs):
    return xs[0] if len(xs) else None


def _and_then(v, f):
    return None if v is None else f(v)


def _step_by(r, n):
    return r[::n]


def _get(xs, *a):
    if isinstance(xs, (list, tuple)):
        (i,) = a
        return xs[i] if 0 <= i < len(xs) else None  # slice::get: None past the end
    return xs.get(*a)  # an object's own get (a sample plane)


def _unwrap_or(v, d):
    return d if v is None else v


def _fill(xs, v):
    xs[:] = [v] * len(xs)


def compile_fn(text, name, extra=None, after=None):
    """A Python function with the parameters and body of Rust fn `name`."""
    params, body = fn_source(text, name, after)
    body = re.sub(r"//[^\n]*", "", body)  # line comments
    names = [p.split(":")[0].strip() for p in _split_top(params, ",") if p.strip()]
    names = [re.sub(r"^&?(mut\s+)?", "", n) for n in names]
    src = "def %s(%s):\n%s\n" % (name, ", ".join(names), "\n".join(block(body)) or "    pass")
    # builtins under names that the reference's locals (`sum`, `min`, ...) cannot shadow
    ns = {"_clamp": _clamp, "_struct": types.SimpleNamespace, "_step_by": _step_by, "_get": _get,
          "_unwrap_or": _unwrap_or, "_first": _first, "_and_then": _and_then, "_RustErr": _RustErr,
          "_unwrap_opt": _unwrap_opt, "_rev": _rev, "_take": _take, "_signum": _signum, "_copy_into": _copy_into, "_deref": _deref, "_USIZE_MAX": _USIZE_MAX, "_abs": abs, "_min": min, "_max": max, "_zip": zip,
          "_enumerate": enumerate, "_list": _clone, "_len": len, "_sum": sum, "_range": range, "_fill": _fill,
          "p_at": lambda line, i: line[7 - i], "q_at": lambda line, j: line[8 + j]}
    ns.update(extra or {})
    exec(compile(src, "<rust:%s>" % name, "exec"), ns)
    fn = ns[name]
    fn.py_source = src
    return fn
