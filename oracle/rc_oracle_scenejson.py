"""CPU oracle of the scene-file reader: SceneSerializer.load and its helpers
(src/main/java/com/retrocam/scene/SceneSerializer.java:72-212) and
Float.parseFloat's grammar and rounding.

TEST INFRASTRUCTURE ONLY: used by tests/ to check librc_host.so's
rcs_json_parse. Nothing in the product imports it.

Values are exact: a decimal or hexadecimal string is turned into a Fraction
and rounded once to the nearest binary32 (ties to even), as Float.parseFloat
rounds; nothing goes through a binary64 intermediate.
"""
import re
from fractions import Fraction

import numpy as np

F = np.float32
_MAXF = Fraction(2) ** 128 - Fraction(2) ** 103  # halfway between FLT_MAX and 2^128 (ties go up: odd)


def round_f32(q: Fraction):
    """The binary32 nearest to q, ties to even (overflow to infinity)."""
    if q == 0:
        return F(0.0)
    neg = q < 0
    a = -q if neg else q
    if a >= _MAXF:
        return F(-np.inf) if neg else F(np.inf)
    # binade: 2^e <= a < 2^(e+1); float32 spacing 2^(e-23), subnormals 2^-149
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    ulp = Fraction(2) ** max(e - 23, -149)
    k = a / ulp
    n = k.numerator // k.denominator
    rem = k - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    v = F(float(n * ulp))  # n * ulp is a binary32 value, exact in binary64
    return -v if neg else v


_DEC = re.compile(r"^(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?$")
_HEX = re.compile(r"^0[xX]([0-9a-fA-F]*)(?:\.([0-9a-fA-F]*))?[pP]([+-]?\d+)$")


def parse_float(s):
    """Float.parseFloat: the value, or None where Java throws NumberFormatException."""
    t = s.strip(" \t\n\r\x0b\x0c" + "".join(chr(c) for c in range(0, 33)))
    if not t:
        return None
    neg = t[0] == "-"
    if t[0] in "+-":
        t = t[1:]
    if t == "NaN":
        return F(np.nan)
    if t == "Infinity":
        return F(-np.inf) if neg else F(np.inf)
    if t and t[-1] in "fFdD":
        t = t[:-1]
    m = _HEX.match(t)
    if m:
        ip, fp, ex = m.group(1) or "", m.group(2) or "", int(m.group(3))
        if not ip and not fp:
            return None
        q = Fraction(int(ip + fp, 16), 16 ** len(fp)) * Fraction(2) ** ex
    else:
        m = _DEC.match(t)
        if not m:
            return None
        q = Fraction(m.group(1) + (m.group(2) or ""))
    v = round_f32(q)
    return F(-v) if neg else v


def _array_body(js, key):
    tok = '"' + key + '"'
    ki = js.find(tok)
    if ki < 0:
        return None
    start = js.find("[", ki + len(tok))
    if start < 0:
        return None
    depth = 0
    for i in range(start, len(js)):
        if js[i] == "[":
            depth += 1
        elif js[i] == "]":
            depth -= 1
            if depth == 0:
                return js[start + 1:i]
    return None


def _split_objects(body):
    out, depth, start = [], 0, -1
    for i, c in enumerate(body):
        if c == "{":
            if depth == 0:
                start = i
            depth += 1
        elif c == "}":
            depth -= 1
            if depth == 0 and start >= 0:
                out.append(body[start:i + 1])
                start = -1
    return out


class JavaThrows(Exception):
    pass


def _str_field(obj, key):
    tok = '"' + key + '"'
    ki = obj.find(tok)
    if ki < 0:
        return None
    scan = ki + len(tok)
    while scan < len(obj) and obj[scan] != '"':
        scan += 1
    if scan >= len(obj):
        return None
    close = scan + 1
    while close < len(obj):
        c = obj[close]
        if c == "\\":
            close += 2
            continue
        if c == '"':
            break
        close += 1
    if close > len(obj):
        raise JavaThrows("substring end past the text")
    return obj[scan + 1:close].replace('\\"', '"').replace("\\\\", "\\")


def _float_field(obj, key):
    tok = '"' + key + '"'
    ki = obj.find(tok)
    if ki < 0:
        return F(0.0)
    i = ki + len(tok)
    while i < len(obj) and obj[i] in ": \t\n\r":
        i += 1
    if i >= len(obj):
        return F(0.0)
    end = i
    while end < len(obj) and (obj[end] in "0123456789.-+eE"):
        end += 1
    v = parse_float(obj[i:end]) if end > i else None
    return F(0.0) if v is None else v


def _java_split_comma(s):
    parts = s.split(",")
    if s:
        while parts and parts[-1] == "":
            parts.pop()
    return parts


def _float_arr(obj, key):
    tok = '"' + key + '"'
    ki = obj.find(tok)
    if ki < 0:
        return None
    start = obj.find("[", ki + len(tok))
    end = obj.find("]", start if start > 0 else 0)
    if start < 0 or end < 0:
        return None
    out = []
    for p in _java_split_comma(obj[start + 1:end]):
        v = parse_float(p.strip())
        out.append(F(0.0) if v is None else v)
    return out


def _f2i(v):
    v = float(v)
    if v != v:
        return 0
    if v >= 2147483647.0:
        return 2147483647
    if v <= -2147483648.0:
        return -2147483648
    return int(v)


def load(text):
    """SceneSerializer.load: (materials, objects), or raises JavaThrows.
    A material is a dict of the Material fields plus its name; an object a dict
    of the SceneObject fields."""
    mats, objs = [], []
    body = _array_body(text, "materials")
    if body is not None:
        for o in _split_objects(body):
            name = _str_field(o, "name")
            alb, em = _float_arr(o, "albedo"), _float_arr(o, "emission")
            m = {"name": name if name is not None else "Material",
                 "albedo": alb if alb is not None and len(alb) == 3 else [F(0.8)] * 3,
                 "emission": em if em is not None and len(em) == 3 else [F(0.0)] * 3,
                 "metallic": _float_field(o, "metallic"),
                 "emissionStrength": _float_field(o, "emissionStrength"),
                 "roughness": _float_field(o, "roughness"),
                 "ior": _float_field(o, "ior")}
            if m["ior"] < F(0.01):
                m["ior"] = F(1.5)
            m["transmission"] = _float_field(o, "transmission")
            mats.append(m)
    body = _array_body(text, "objects")
    if body is not None:
        for o in _split_objects(body):
            name = _str_field(o, "name")
            typ = _str_field(o, "type")
            d = {k: _float_field(o, k) for k in ("px", "py", "pz", "sx", "sy", "sz", "rx", "ry", "rz")}
            stacks = _f2i(_float_field(o, "stacks"))
            slices = _f2i(_float_field(o, "slices"))
            mi = _f2i(_float_field(o, "materialIndex"))
            d["name"] = name if name is not None else "Object"
            d["type"] = "SPHERE" if typ is not None and typ.upper() == "SPHERE" else "BOX"
            d["materialIndex"] = max(0, mi)
            d["stacks"] = stacks if stacks > 2 else 24
            d["slices"] = slices if slices > 2 else 24
            objs.append(d)
    return mats, objs
