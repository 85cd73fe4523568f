"""serde_json ``to_string_pretty`` emulation (oracle for the library's JSON writer).

Follows the output conventions the reference relies on (main.rs:893, lib.rs:31-45):
2-space indentation, ``"key": value``, empty arrays as ``[]``, ``null`` for None and for
non-finite floats, integers bare, and f64 through the ryu shortest round-trip digits laid
out as: integral values with ``.0`` up to 16 digits before the point, plain decimals for
exponents -5 < k <= 16, otherwise ``d.ddde<exp>`` (``1e-7``, ``1.5e20``; no '+', no
zero padding).
"""
import math
from decimal import Decimal


def fmt_f64(x):
    x = float(x)
    if not math.isfinite(x):
        return "null"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    d = Decimal(repr(abs(x))).normalize()
    _, digits, exp = d.as_tuple()
    ds = "".join(str(v) for v in digits)
    length = len(ds)
    kk = length + exp  # 10^(kk-1) <= v < 10^kk
    if 0 <= exp and kk <= 16:
        body = ds + "0" * exp + ".0"
    elif 0 < kk <= 16:
        body = ds[:kk] + "." + ds[kk:]
    elif -5 < kk <= 0:
        body = "0." + "0" * (-kk) + ds
    elif length == 1:
        body = ds + "e" + str(kk - 1)
    else:
        body = ds[0] + "." + ds[1:] + "e" + str(kk - 1)
    return sign + body


def dumps_pretty(v, indent=0):
    pad = "  " * indent
    pad1 = "  " * (indent + 1)
    if v is None:
        return "null"
    if v is True:
        return "true"
    if v is False:
        return "false"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        return fmt_f64(v)
    if isinstance(v, str):
        out = ['"']
        for ch in v:
            if ch == '"':
                out.append('\\"')
            elif ch == "\\":
                out.append("\\\\")
            elif ch == "\n":
                out.append("\\n")
            elif ord(ch) < 0x20:
                out.append("\\u%04x" % ord(ch))
            else:
                out.append(ch)
        out.append('"')
        return "".join(out)
    if isinstance(v, (list, tuple)):
        if not v:
            return "[]"
        return "[\n" + ",\n".join(pad1 + dumps_pretty(e, indent + 1) for e in v) + "\n" + pad + "]"
    if isinstance(v, dict):
        if not v:
            return "{}"
        return "{\n" + ",\n".join(pad1 + dumps_pretty(k, indent + 1) + ": " +
                                  dumps_pretty(e, indent + 1) for k, e in v.items()) + "\n" + pad + "}"
    raise TypeError(type(v))
