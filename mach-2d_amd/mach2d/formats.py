"""On-disk formats around the SIMPLEC path.

ClassIfile mirrors class_ifile (src/util/class_ifile.f90), the reader of
MACH-2D's parameter files: every line "value FS varname FS description"
(FS a set of separator characters, "&" in the shipped files) defines a
variable; other lines are comments.  get_value(name, kind) returns the value
read as an int, a float or a string (the Fortran generic get_value picks the
reading by the type of its output argument); an unknown name stops the
program, as the reference does (SystemExit here).  solvers.solvers_init
accepts a ClassIfile.

residual_header / residual_row give the lines of the residual file written
by mach2d_main.f90:495-517: format (A10, 4(1X, A14), A) for the header and
(I10, 4(1X, ES14.7), A) for a row (it, dt, norm, max|pl|/p_avg, rmass).
"""
from __future__ import annotations

import math
import sys

CHLEN = 500            # character length of a stored field (class_ifile.f90:62)
LINE = 3 * CHLEN       # length of the line buffer (:146)


def _next_token(s: str, fs: str):
    """next_token (:258-298) on the padded line buffer: (token, rest, last)."""
    idx = next((k for k, ch in enumerate(s) if ch in fs), -1)
    if idx < 0:
        return s, "", True
    tok = s[:idx]
    if idx + 1 < len(s):
        return tok, s[idx + 1:], False
    return tok, "", True


class ClassIfile:
    """class_ifile: init(filename, field_separator), load(), get_value,
    set_value, print_read_data, print_pickedup_data."""

    def init(self, filename, field_separator):
        self.filename = filename
        self.fs = field_separator.rstrip() or field_separator
        self.data = []        # [value, varname, description]
        self.pickedup = []
        return self

    def clear(self):
        self.data = []
        self.pickedup = []

    def load(self):
        """load (:139-203): a line of three or more fields adds an entry; the
        third field takes the rest of the line."""
        self.clear()
        with open(self.filename, "r", encoding="latin-1") as fh:
            for raw in fh:
                line = raw.rstrip("\n").rstrip("\r")[:LINE].ljust(LINE)
                toks = []
                rest = line
                while True:
                    tok, rest, last = _next_token(rest, self.fs)
                    toks.append(tok)
                    if not last and len(toks) == 2:
                        toks.append(rest.strip())
                        break
                    if last or len(toks) == 3:
                        break
                if len(toks) == 3:
                    self.data.append([toks[0].rstrip()[:CHLEN], toks[1].rstrip()[:CHLEN],
                                      toks[2].rstrip()[:CHLEN]])
        names = [d[1].strip() for d in self.data]
        for i, n in enumerate(names):   # check_var_repetition (:207-229)
            if n in names[i + 1:]:
                print("%s: '%s' was defined repeatedly." % (self.filename.strip(), n))
        return self

    def _find(self, varname):
        key = varname.strip()
        for i, d in enumerate(self.data):
            if d[1].strip() == key:
                return i
        print(" Variable %s not found. Stopping." % key)
        raise SystemExit(" Variable %s not found. Stopping." % key)

    def set_value(self, value, varname):
        self.data[self._find(varname)][0] = str(value)

    def get_value(self, varname, kind=str):
        """The value of `varname` read as kind (int, float or str), as the
        generic get_value reads it: list-directed for numbers (the first item
        of the field; a Fortran exponent letter d/D/q is accepted), the whole
        field left- and right-trimmed for strings."""
        i = self._find(varname)
        if i not in self.pickedup:
            self.pickedup.append(i)
        v = self.data[i][0]
        if kind is str:
            return v.strip()
        item = v.replace(",", " ").split()[0] if v.split() else ""
        if kind is int:
            return int(item)
        if kind is float:
            return float(item.replace("d", "e").replace("D", "e").replace("q", "e")
                         .replace("Q", "e"))
        raise TypeError("kind must be int, float or str")

    def print_read_data(self, unt=None, frm=None):
        out = unt or sys.stdout
        for v, n, d in self.data:
            if frm:
                out.write(frm % (v.strip(), n.strip(), d.strip()) + "\n")
            else:   # (2(1X,A30),2X,A)
                out.write(" %30s %30s  %s\n" % ((v.strip() + " &")[:30].ljust(30),
                                                (n.strip() + " &")[:30].ljust(30), d.strip()))

    def print_pickedup_data(self, unt=None):
        out = unt or sys.stdout
        for i in self.pickedup:
            v, n, d = self.data[i]
            out.write(" %30s %30s  %s\n" % ((v.strip() + " &")[:30].ljust(30),
                                            (n.strip() + " &")[:30].ljust(30), d.strip()))


def es14_7(x: float) -> str:
    """Fortran ES14.7: d.ddddddd E+ee right-aligned in 14 characters (a
    three-digit exponent drops the letter, E+eee -> +eee)."""
    if math.isnan(x):
        return "NaN".rjust(14)
    if math.isinf(x):
        return ("-Infinity" if x < 0 else "Infinity").rjust(14)
    m, e = ("%.7E" % x).split("E")
    ei = int(e)
    ex = ("E%+03d" % ei) if abs(ei) < 100 else ("%+04d" % ei)
    s = m + ex
    return s.rjust(14) if len(s) <= 14 else "*" * 14


def residual_header(msg: str = "") -> str:
    """(A10, 4(1X, A14), A) with it, dt, norm, max(|pl|)/p_avg, rmass (an A
    edit keeps the leftmost w characters of a longer string)."""
    cols = ["it".rjust(10)] + [" " + c[:14].rjust(14) for c in
                               ("dt", "norm", "max(|pl|)/p_avg", "rmass")]
    return "".join(cols) + msg


def residual_row(it: int, dt: float, norm: float, normpl: float, rmass: float,
                 msg: str = "") -> str:
    """(I10, 4(1X, ES14.7), A)."""
    i = str(int(it))
    head = i.rjust(10) if len(i) <= 10 else "*" * 10
    return head + "".join(" " + es14_7(float(v)) for v in (dt, norm, normpl, rmass)) + msg.rstrip()
