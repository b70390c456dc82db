"""Reader for laboetie's `lb.in` input file, same grammar as the reference.

The reference parser (module_input.f90) reads the file into 100-character
lines, left-adjusts and trims them, blanks everything from a `#` on, drops
blank lines (put_input_in_character_array, module_input.f90:561-633), and
looks a tag up as "the first line whose first len(tag) characters are the tag
and whose next character is a blank"; the value is read list-directed from
column len(tag)+4 to len(tag)+50 (input_dp and siblings, :62-75).  So
`tag = value` with one blank either side of `=` is the accepted form, the first
match wins, and tags are case-sensitive.  Defaults live at the call sites.
"""
import re

LINE_LEN = 100


class LbInError(ValueError):
    pass


def _fortran_lines(text):
    # read '(a)' into character(len=100) keeps the first 100 characters;
    # adjustl / trim shift blanks only (a tab stays a tab, so a tag followed
    # by a tab is not found, as in the reference)
    lines = []
    for raw in text.splitlines():
        line = raw[:LINE_LEN].lstrip(" ").rstrip(" ")
        k = line.find("#")
        if k >= 0:
            line = line[:k]
        line = line.strip(" ")
        if line:
            lines.append(line.ljust(LINE_LEN))
    return lines


def _list_directed_tokens(field):
    # list-directed input: values separated by blanks or commas, '/' ends the record
    field = field.split("/")[0]
    return [t for t in re.split(r"[\s,]+", field.strip()) if t]


def _to_float(tok):
    return float(tok.replace("d", "e").replace("D", "e"))


def _to_logical(tok):
    t = tok.lstrip(".")
    if not t:
        raise LbInError(f"bad logical {tok!r}")
    c = t[0].upper()
    if c == "T":
        return True
    if c == "F":
        return False
    raise LbInError(f"bad logical {tok!r}")


class LbIn:
    """Parsed lb.in; getters mirror getinput%{int,dp,dp3,log,char}."""

    def __init__(self, text):
        self.lines = _fortran_lines(text)

    @classmethod
    def from_file(cls, path):
        with open(path) as fh:
            return cls(fh.read())

    def _field(self, tag):
        j = len(tag)
        for line in self.lines:
            if line[:j] == tag and line[j:j + 1] == " ":
                return line[j + 3:j + 50]
        return None

    def _tokens(self, tag, count, default, conv):
        f = self._field(tag)
        if f is None:
            if default is None:
                raise LbInError(f"tag {tag!r} not found in lb.in and no default")
            return default
        toks = _list_directed_tokens(f)
        if len(toks) < count:
            raise LbInError(f"tag {tag!r}: expected {count} value(s), got {f.strip()!r}")
        vals = [conv(t) for t in toks[:count]]
        return vals[0] if count == 1 else vals

    def has(self, tag):
        return self._field(tag) is not None

    def int(self, tag, default=None, assert_=None):
        v = self._tokens(tag, 1, default, lambda t: int(_to_float(t)) if ("." in t or "e" in t.lower()) else int(t))
        _check(tag, v, assert_)
        return v

    def dp(self, tag, default=None, assert_=None):
        v = self._tokens(tag, 1, default, _to_float)
        _check(tag, v, assert_)
        return float(v)

    def dp3(self, tag, default=None):
        return [float(x) for x in self._tokens(tag, 3, default, _to_float)]

    def int3(self, tag, default=None):
        return [int(x) for x in self._tokens(tag, 3, default, int)]

    def log(self, tag, default=None):
        return bool(self._tokens(tag, 1, default, _to_logical))

    def char(self, tag, default=None):
        f = self._field(tag)
        if f is None:
            if default is None:
                raise LbInError(f"tag {tag!r} not found in lb.in and no default")
            return default
        toks = _list_directed_tokens(f)
        return toks[0].strip("'\"") if toks else ""


def _check(tag, v, assert_):
    if assert_ is None:
        return
    ok = {">0": v > 0, ">=0": v >= 0, "<0": v < 0, "<=0": v <= 0}.get(assert_)
    if ok is None:
        raise LbInError(f"unknown assert {assert_!r} for tag {tag!r}")
    if not ok:
        raise LbInError(f"{tag} = {v}. Must be {assert_}")


def read_geom_in(text):
    """geom.in: one `i j k` (1-based) solid node per line
    (construct_custom, module_geometry.f90:380-426).  Returns an (n, 3) int32
    array; the reference re-reads the last line at end of file, which marks
    the same node twice and changes nothing."""
    import numpy as np

    pts = []
    for ln in text.splitlines():
        toks = _list_directed_tokens(ln)
        if len(toks) >= 3:
            pts.append([int(toks[0]), int(toks[1]), int(toks[2])])
    return np.array(pts, dtype=np.int32).reshape(-1, 3)
