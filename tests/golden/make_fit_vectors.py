"""Generate tests/golden/fit_vectors.npz from the reference's own source text.

Run in a container where the reference is mounted (read-only, text only):
    python tests/golden/make_fit_vectors.py /root/reference

Independent of the oracle's hand-written C: this script parses the parameter
declarations of props.F90 and the coefficient declarations and return
expressions of each fit function in ptcl_props.F90, and evaluates them with a
small expression evaluator (no Python eval), widening default-real literals
through float32 as the Fortran compiler does.  Only data (T grid and fit
values) is written; no reference text is stored.
"""
from __future__ import annotations

import os
import re
import sys

import numpy as np

FUNCS = ["get_mue", "get_mut", "get_De", "get_Dt", "get_k_ex", "get_k_ir", "get_nu",
         "get_k_sc", "get_k_si", "get_beta"]

NUM = re.compile(r"(\d+\.\d*|\.\d+|\d+)([eEdD][-+]?\d+)?")


def lit(tok: str) -> float:
    """Fortran literal -> value: 'd' exponent is double, otherwise float32-widened."""
    if "d" in tok.lower():
        return float(tok.lower().replace("d", "e"))
    if re.fullmatch(r"\d+", tok):
        return float(int(tok))  # integer literal
    return float(np.float32(float(tok)))


def tokenize(s: str):
    s = s.replace("&", " ")
    i, out = 0, []
    while i < len(s):
        c = s[i]
        if c.isspace():
            i += 1
        elif s.startswith("**", i):
            out.append(("op", "**"))
            i += 2
        elif c in "+-*/(),<>":
            out.append(("op", c))
            i += 1
        elif c.isdigit() or (c == "." and i + 1 < len(s) and s[i + 1].isdigit()):
            m = NUM.match(s, i)
            out.append(("num", m.group(0)))
            i = m.end()
        elif c.isalpha() or c == "_":
            m = re.match(r"[A-Za-z_][A-Za-z_0-9%]*", s[i:])
            out.append(("id", m.group(0)))
            i += m.end()
        else:
            raise ValueError(f"unexpected {c!r} in {s!r}")
    return out


class Expr:
    """Recursive-descent evaluator for the Fortran arithmetic in the fits."""

    def __init__(self, toks, env, fenv):
        self.t, self.i, self.env, self.fenv = toks, 0, env, fenv

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self):
        tok = self.t[self.i]
        self.i += 1
        return tok

    def parse(self):
        v = self.add()
        if self.i != len(self.t):
            raise ValueError(f"trailing tokens {self.t[self.i:]}")
        return v

    def add(self):
        k, v0 = self.peek()
        if k == "op" and v0 in "+-":
            self.take()
            v = self.mul()
            v = -v if v0 == "-" else v
        else:
            v = self.mul()
        while True:
            k, op = self.peek()
            if k == "op" and op in ("+", "-"):
                self.take()
                r = self.mul()
                v = v + r if op == "+" else v - r
            else:
                return v

    def mul(self):
        v = self.power()
        while True:
            k, op = self.peek()
            if k == "op" and op in ("*", "/"):
                self.take()
                r = self.power()
                v = v * r if op == "*" else v / r
            else:
                return v

    def power(self):
        base = self.atom()
        k, op = self.peek()
        if k == "op" and op == "**":
            self.take()
            kk, tok = self.peek()
            if kk == "num":
                self.take()
                if re.fullmatch(r"\d+", tok):  # integer exponent: repeated product
                    n = int(tok)
                    out = np.ones_like(base) if isinstance(base, np.ndarray) else 1.0
                    for _ in range(n):
                        out = out * base
                    return out
                return np.power(base, lit(tok))
            return np.power(base, self.atom())
        return base

    def atom(self):
        k, tok = self.take()
        if k == "num":
            return lit(tok)
        if k == "op" and tok == "(":
            v = self.add()
            self.take()
            return v
        if k == "id":
            name = tok.lower()
            kk, nxt = self.peek()
            if kk == "op" and nxt == "(" and name in self.fenv:
                self.take()
                args = [self.add()]
                while self.peek() == ("op", ","):
                    self.take()
                    args.append(self.add())
                self.take()
                return self.fenv[name](*args)
            return self.env[name]
        raise ValueError(f"bad token {tok!r}")


def f32_log(x):
    return np.float64(np.float32(np.log(np.float64(np.float32(x)))))


FENV = {"log": np.log, "exp": np.exp, "log10": np.log10, "sqrt": np.sqrt,
        "max": np.maximum, "min": np.minimum, "atan": np.arctan}


def statements(text: str):
    """Join Fortran continuation lines and drop comments."""
    out, cur = [], ""
    for line in text.splitlines():
        line = line.split("!")[0].rstrip()
        if not line.strip():
            continue
        if line.rstrip().endswith("&"):
            cur += line.rstrip()[:-1] + " "
            continue
        out.append((cur + line).strip())
        cur = ""
    return out


def props_env(ref: str) -> dict:
    env = {}
    for st in statements(open(os.path.join(ref, "props.F90")).read()):
        m = re.match(r"real\(8\),\s*parameter\s*::\s*(.*)", st, re.I)
        if not m:
            continue
        for part in re.split(r",\s*(?=[A-Za-z_]\w*\s*=)", m.group(1)):
            name, expr = part.split("=", 1)
            env[name.strip().lower()] = float(Expr(tokenize(expr), env, FENV).parse())
    return env


def fit_funcs(ref: str, env: dict):
    text = open(os.path.join(ref, "ptcl_props.F90")).read()
    fns = {}
    for name in FUNCS:
        m = re.search(rf"function\s+{name}\s*\(T\)(.*?)end function", text, re.S | re.I)
        body = statements(m.group(1))
        loc = {}
        xexpr, branches, rexpr, cond = None, [], None, None
        for st in body:
            dm = re.match(r"real\(8\)\s*::\s*(.*)", st, re.I)
            if dm:
                for dec in re.split(r",\s*(?=[A-Za-z_]\w*\s*=)", dm.group(1)):
                    if "=" in dec:
                        nm, v = dec.split("=", 1)
                        nm = nm.split(",")[-1].strip().lower()
                        loc[nm] = lit(v.strip())
                continue
            if re.match(r"x\s*=", st, re.I):
                xexpr = st.split("=", 1)[1]
            elif re.match(rf"{name}\s*=", st, re.I):
                branches.append(st.split("=", 1)[1])
            elif st.lower().startswith("if"):
                cond = re.search(r"log\(([\d.]+)\)", st).group(1)
        fns[name] = (xexpr, branches, cond, loc)
    return fns


def evaluate(ref: str, T: np.ndarray) -> dict:
    env = props_env(ref)
    fns = fit_funcs(ref, env)
    out = {}
    for name, (xexpr, branches, cond, loc) in fns.items():
        e = dict(env)
        e.update(loc)
        e["t"] = T
        x = Expr(tokenize(xexpr), e, FENV).parse()
        e["x"] = x
        vals = [Expr(tokenize(b), e, dict(FENV, **{"10d0": None})).parse() for b in branches]
        if cond is not None:
            out[name] = np.where(x < f32_log(float(cond)), vals[0], vals[1])
        else:
            out[name] = vals[0]
    out["_ph0"] = np.array(env["ph0"])
    return out


def ptcl_params(ref: str, env: dict) -> dict:
    """Parameter constants of ptcl_props.F90 (module header), evaluated like props."""
    text = open(os.path.join(ref, "ptcl_props.F90")).read()
    head = text[: text.lower().index("contains")]
    out = dict(env)
    for st in statements(head):
        m = re.match(r"real\(8\),\s*parameter\s*::\s*(.*)", st, re.I)
        if not m:
            continue
        for part in re.split(r",\s*(?=[A-Za-z_]\w*\s*=)", m.group(1)):
            name, expr = part.split("=", 1)
            out[name.strip().lower()] = float(Expr(tokenize(expr), out, FENV).parse())
    return out


def python_restatements(ref: str, x: np.ndarray) -> dict:
    """The reference's own Python restatements with double literals:
    coeftest.py getDe/getDt (its x0/ninf*t0 scale is commented out) and
    wvgRes_data.py nu (the argument of np.exp only), evaluated at the same x."""
    out = {}
    for fname, funcs in (("coeftest.py", ("getDt", "getDe")), ("wvgRes_data.py", ("nu",))):
        text = open(os.path.join(ref, fname)).read()
        for fn in funcs:
            m = re.search(rf"def {fn}\(\w+\):(.*?)return", text, re.S)
            body = m.group(1)
            coef = {k: float(v) for k, v in re.findall(r"^\s*(\w)\s*=\s*([-+\d.eE]+)\s*$",
                                                         body, re.M)}
            rest = text[m.end():]
            i0 = rest.index("np.exp(") + len("np.exp(")
            depth, i = 1, i0
            while depth:
                depth += {"(": 1, ")": -1}.get(rest[i], 0)
                i += 1
            arg = rest[i0:i - 1].replace("\n", " ")
            var = "y" if re.search(r"\by\b", arg) else "x"
            env = dict(coef)
            env[var] = x
            toks = [(k, v) if k != "num" else ("num", v + "d0" if "e" not in v.lower()
                                                 else v.lower().replace("e", "d"))
                    for k, v in tokenize(arg)]
            out[f"{fname}:{fn}"] = np.exp(Expr(toks, env, FENV).parse())
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    env = props_env(ref)
    penv = ptcl_params(ref, env)
    v = np.concatenate([np.geomspace(0.05, 400.0, 401),
                        [0.234, 0.9611, 1.37, 0.5, 16.0, 157.0, 5.119, 5.667, 7.5]])
    T = v / env["ph0"]
    res = evaluate(ref, T)
    xa = np.log(np.maximum(0.234, np.minimum(157.0, T * env["ph0"])))
    py = python_restatements(ref, xa)
    consts = {f"c_{k}": np.array(penv[k]) for k in
              ["eps0", "mu0", "e", "kb", "me", "x0", "ph0", "t0", "wp02", "w2", "tg", "p",
               "ninf", "n_zero", "mi", "mui", "vi", "di", "dm", "k_r", "k_mp", "k_2q", "k_3q",
               "h_ir", "h_ex", "h_si", "h_sc", "gam"]}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fit_vectors.npz")
    np.savez(dst, T=T, p_torr=np.array(env["p"]), **{k: res[k] for k in FUNCS},
             py_De=py["coeftest.py:getDe"], py_Dt=py["coeftest.py:getDt"],
             py_nu=py["wvgRes_data.py:nu"], **consts)
    print(f"wrote {dst}: {len(T)} points, p = {env['p']} Torr, ph0 = {env['ph0']!r}")


if __name__ == "__main__":
    main()
