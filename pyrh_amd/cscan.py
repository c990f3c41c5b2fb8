"""Minimal emulation of C ``sscanf`` / ``fgets`` line handling.

RH's readers parse fixed-column records with ``sscanf`` at character
offsets (e.g. rh/kurucz.c:178-330).  ``sscanf`` differs from
``str.split`` when fields touch (``"630.1500-10.123"`` is two numbers) and
in how it reports an exhausted input (EOF, -1).  Only the conversions used
by the readers are supported: ``%d``, ``%lf``/``%f``, ``%s``, ``%<w>d``,
``%<w>s``, ``%<w>c`` and literal characters.
"""
from __future__ import annotations

import re

_WS = " \t\n\r\v\f"
_FLOAT_RE = re.compile(
    r"[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan)",
    re.IGNORECASE)
_INT_RE = re.compile(r"[+-]?\d+")
_SPEC_RE = re.compile(r"%(\d*)(l?)([dfsc])")


def sscanf(text: str, fmt: str):
    """Return (count, values) like C sscanf (count is -1 on input failure
    before the first conversion)."""
    vals = []
    pos = 0
    fpos = 0
    n = len(text)
    nconv = 0
    while fpos < len(fmt):
        ch = fmt[fpos]
        if ch in _WS:
            while pos < n and text[pos] in _WS:
                pos += 1
            while fpos < len(fmt) and fmt[fpos] in _WS:
                fpos += 1
            continue
        if ch != "%":
            if pos >= n:
                return (-1 if nconv == 0 else nconv), vals
            if text[pos] != ch:
                return nconv, vals
            pos += 1
            fpos += 1
            continue
        m = _SPEC_RE.match(fmt, fpos)
        if not m:
            raise ValueError(f"unsupported format {fmt!r}")
        width = int(m.group(1)) if m.group(1) else None
        conv = m.group(3)
        fpos = m.end()
        if conv != "c":
            while pos < n and text[pos] in _WS:
                pos += 1
        if pos >= n:
            return (-1 if nconv == 0 else nconv), vals
        seg = text[pos:pos + width] if width else text[pos:]
        if conv == "c":
            w = width or 1
            if n - pos < w:
                return (-1 if nconv == 0 else nconv), vals
            vals.append(text[pos:pos + w])
            pos += w
        elif conv == "s":
            j = 0
            while j < len(seg) and seg[j] not in _WS:
                j += 1
            vals.append(seg[:j])
            pos += j
        elif conv == "d":
            mm = _INT_RE.match(seg)
            if not mm:
                return nconv, vals
            vals.append(int(mm.group(0)))
            pos += mm.end()
        else:
            mm = _FLOAT_RE.match(seg)
            if not mm:
                return nconv, vals
            vals.append(float(mm.group(0)))
            pos += mm.end()
        nconv += 1
    return nconv, vals


def data_lines(path, comment: str = "#"):
    """Lines as RH's getLine() returns them (rh/getline.c:30-50): skip
    lines that are blank or whose first character is the comment
    character.  Lines keep their trailing newline."""
    with open(path, "r", errors="replace") as f:
        for line in f:
            if line.strip() and line[0] != comment:
                yield line
