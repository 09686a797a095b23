"""Reader of OpenFOAM dictionary text (the solidDict format the reference reads with
``Foam::dictionary``, e.g. src/solidcloud.cpp:43-160 and tool_vof/solidcloud.cpp).

Grammar handled: ``key value ... ;`` entries, ``key { ... }`` sub-dictionaries, parenthesised
lists ``( a b c )`` (nested), words and numbers, ``//`` and ``/* */`` comments.  The
``FoamFile`` header is returned like any sub-dictionary.  Macros, ``#include`` and regular
expression keys are not supported (the coupling step's dictionaries do not use them).
"""
from __future__ import annotations

import re

__all__ = ["parse_foam_dict", "FoamDictError"]


class FoamDictError(ValueError):
    pass


_TOKEN = re.compile(r'\s+|//[^\n]*|/\*.*?\*/|"[^"]*"|[{}();]|[^\s{}();"]+', re.S)


def _tokens(text: str):
    for m in _TOKEN.finditer(text):
        t = m.group(0)
        if t.isspace() or t.startswith("//") or t.startswith("/*"):
            continue
        yield t


def _atom(t: str):
    if t.startswith('"'):
        return t[1:-1]
    try:
        return int(t)
    except ValueError:
        pass
    try:
        return float(t)
    except ValueError:
        return t


def parse_foam_dict(text: str) -> dict:
    toks = list(_tokens(text))
    pos = 0

    def parse_list():
        nonlocal pos
        out = []
        while True:
            if pos >= len(toks):
                raise FoamDictError("unterminated list")
            t = toks[pos]
            pos += 1
            if t == ")":
                return out
            if t == "(":
                out.append(parse_list())
            elif t in "{};":
                raise FoamDictError(f"unexpected {t!r} in a list")
            else:
                out.append(_atom(t))

    def parse_dict(top: bool) -> dict:
        nonlocal pos
        out = {}
        while pos < len(toks):
            key = toks[pos]
            pos += 1
            if key.startswith('"'):
                key = key[1:-1]
            if key == "}":
                if top:
                    raise FoamDictError("unbalanced '}'")
                return out
            if key in "{();":
                if key == ";":
                    continue
                raise FoamDictError(f"expected a keyword, got {key!r}")
            if pos < len(toks) and toks[pos] == "{":
                pos += 1
                out[key] = parse_dict(False)
                continue
            vals = []
            while True:
                if pos >= len(toks):
                    raise FoamDictError(f"entry {key!r} not terminated by ';'")
                t = toks[pos]
                pos += 1
                if t == ";":
                    break
                if t == "(":
                    vals.append(parse_list())
                elif t in "{})":
                    raise FoamDictError(f"unexpected {t!r} in entry {key!r}")
                else:
                    vals.append(_atom(t))
            out[key] = vals[0] if len(vals) == 1 else vals
        if not top:
            raise FoamDictError("unterminated dictionary")
        return out

    return parse_dict(True)
