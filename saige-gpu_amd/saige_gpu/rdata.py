"""R's serialization format (version 3, XDR), for the null-model file that
SAIGE step 2 loads: fitNULLGLMM ends with save(modglmm, file=...) (FG.R:1297,
1536), a gzip'd stream "RDX3\\n" + serialize() of a pairlist that binds each
saved name to its object.

Only data travels: NULL, logical / integer / double / character vectors,
lists, matrices (a dim attribute) and other attributes (names, class).  The
reader also follows the reference table and the ALTREP compact sequences
and wrappers that R writes; it refuses code objects (closures, environments,
promises, bytecode, external pointers).  Nothing read is evaluated.

Python side: dict <-> named list, list/tuple <-> unnamed list, numpy arrays
and scalars <-> atomic vectors (2-D -> matrix, column-major), str <->
character, None <-> NULL, bool <-> logical.  Attributes other than names
and dim ride in RObj(value, attrs)."""
import gzip
import io
import struct

import numpy as np

# SEXP types (R Internals, "Serialization formats")
NILSXP, SYMSXP, LISTSXP, CLOSXP, ENVSXP, PROMSXP, LANGSXP = 0, 1, 2, 3, 4, 5, 6
CHARSXP, LGLSXP, INTSXP, REALSXP, CPLXSXP, STRSXP, VECSXP = 9, 10, 13, 14, 15, 16, 19
EXPRSXP, BCODESXP, EXTPTRSXP, S4SXP = 20, 21, 22, 25
REFSXP, NILVALUE_SXP, GLOBALENV_SXP, EMPTYENV_SXP, BASEENV_SXP = 255, 254, 253, 242, 241
MISSINGARG_SXP, ALTREP_SXP, ATTRLISTSXP = 251, 238, 239

IS_OBJECT, HAS_ATTR, HAS_TAG = 1 << 8, 1 << 9, 1 << 10
UTF8_MASK, ASCII_MASK = 1 << 3, 1 << 6
NA_INTEGER = -(2 ** 31)
R_VERSION = (4 << 16) | (1 << 8) | 2          # written as R 4.1.2
MIN_READER = (3 << 16) | (5 << 8) | 0         # readable by R >= 3.5.0


class RObj:
    """A value with R attributes other than names/dim (e.g. class)."""

    def __init__(self, value, attrs):
        self.value = value
        self.attrs = attrs

    def __repr__(self):
        return f"RObj({self.value!r}, attrs={list(self.attrs)})"


class RFormatError(ValueError):
    pass


# ------------------------------------------------------------------ reader

class _Reader:
    def __init__(self, buf):
        self.b = buf
        self.o = 0
        self.refs = []

    def int(self):
        v = struct.unpack_from(">i", self.b, self.o)[0]
        self.o += 4
        return v

    def length(self):
        n = self.int()
        if n == -1:
            hi, lo = self.int(), self.int()
            n = (hi << 32) + (lo & 0xFFFFFFFF)
        return n

    def raw(self, n):
        v = self.b[self.o:self.o + n]
        if len(v) != n:
            raise RFormatError("truncated stream")
        self.o += n
        return v

    def item(self):
        flags = self.int()
        t = flags & 0xFF
        has_attr, has_tag = bool(flags & HAS_ATTR), bool(flags & HAS_TAG)
        if t == NILVALUE_SXP:
            return None
        if t == REFSXP:
            i = flags >> 8
            if i == 0:
                i = self.int()
            return self.refs[i - 1]
        if t in (GLOBALENV_SXP, EMPTYENV_SXP, BASEENV_SXP, MISSINGARG_SXP):
            raise RFormatError(f"environment/special object (type {t}) in a data file")
        if t == SYMSXP:
            name = self.item()
            sym = _Sym(name)
            self.refs.append(sym)
            return sym
        if t in (LISTSXP, ATTRLISTSXP):
            return self._pairlist(has_attr, has_tag)
        if t == CHARSXP:
            n = self.int()
            if n == -1:
                return None                      # NA_character_
            latin1 = (flags >> 12) & (1 << 2)
            return self.raw(n).decode("latin-1" if latin1 else "utf-8")
        if t in (LGLSXP, INTSXP):
            n = self.length()
            v = np.frombuffer(self.raw(4 * n), ">i4").astype(np.int32)
            if t == LGLSXP:
                v = np.where(v == NA_INTEGER, -1, v).astype(np.int8)   # -1 marks NA
                v = _Lgl(v)
            return self._attrs(v, has_attr)
        if t == REALSXP:
            n = self.length()
            return self._attrs(np.frombuffer(self.raw(8 * n), ">f8").astype(np.float64), has_attr)
        if t == CPLXSXP:
            n = self.length()
            d = np.frombuffer(self.raw(16 * n), ">f8").astype(np.float64)
            return self._attrs(d[0::2] + 1j * d[1::2], has_attr)
        if t == STRSXP:
            n = self.length()
            return self._attrs(RStr(self.item() for _ in range(n)), has_attr)
        if t in (VECSXP, EXPRSXP):
            n = self.length()
            return self._attrs([self.item() for _ in range(n)], has_attr, vector=True)
        if t == ALTREP_SXP:
            info = self.item()
            state = self.item()
            attr = self.item()
            return self._altrep(info, state, attr)
        raise RFormatError(f"unsupported SEXP type {t} (code objects are not read)")

    def _pairlist(self, has_attr, has_tag):
        attr = self.item() if has_attr else None
        tag = self.item() if has_tag else None
        car = self.item()
        cdr = self.item()
        node = [(tag.name if isinstance(tag, _Sym) else None, car)]
        tail = None
        if isinstance(cdr, _Pairlist):
            node.extend(cdr.items)
            tail = cdr.tail
        elif cdr is not None:
            tail = cdr                        # dotted pair (e.g. an ALTREP state)
        pl = _Pairlist(node)
        pl.tail = tail
        if attr is not None:
            pl.attr = attr
        return pl

    def _attrs(self, v, has_attr, vector=False):
        if not has_attr:
            return v
        at = self.item()
        attrs = {k: a for k, a in at.items} if isinstance(at, _Pairlist) else {}
        return _apply_attrs(v, attrs)

    def _altrep(self, info, state, attr):
        cls = info.items[0][1].name if isinstance(info, _Pairlist) else None
        if cls == "compact_intseq":
            n, start, step = (int(x) for x in state)
            v = (start + step * np.arange(n)).astype(np.int32)
        elif cls == "compact_realseq":
            n, start, step = (float(x) for x in state)
            v = start + step * np.arange(int(n), dtype=np.float64)
        elif cls in ("wrap_integer", "wrap_real", "wrap_logical", "wrap_string", "wrap_complex", "wrap_list"):
            v = state[0]
        elif cls == "deferred_string":
            arg = state.items[0][1] if isinstance(state, _Pairlist) else state[0]
            v = RStr(("%d" % x) if isinstance(x, (int, np.integer)) else repr(float(x)) for x in np.asarray(arg))
        else:
            raise RFormatError(f"unsupported ALTREP class {cls!r}")
        if isinstance(attr, _Pairlist):
            return _apply_attrs(v, {k: a for k, a in attr.items})
        return v


class RStr(list):
    """A character vector (None for NA)."""


class _Sym:
    __slots__ = ("name",)

    def __init__(self, name):
        self.name = name


class _Pairlist:
    def __init__(self, items):
        self.items = items
        self.attr = None
        self.tail = None


class _Lgl(np.ndarray):
    """Logical vector marker (int8, 1 TRUE, 0 FALSE, -1 NA)."""

    def __new__(cls, a):
        return np.asarray(a, np.int8).view(cls)


def _apply_attrs(v, attrs):
    attrs = dict(attrs)
    names = attrs.pop("names", None)
    dim = attrs.pop("dim", None)
    if dim is not None and isinstance(v, np.ndarray):
        v = np.asarray(v).reshape(tuple(int(d) for d in dim), order="F")
    if (names is not None and isinstance(v, list) and not isinstance(v, RStr) and len(names) == len(v)
            and all(n is not None for n in names)):
        v = dict(zip(names, v))
    elif names is not None:
        attrs["names"] = names
    return RObj(v, attrs) if attrs else v


def loads(data):
    """Parse an .rda body (gzip'd or not) into {name: object}."""
    if data[:2] == b"\x1f\x8b":
        data = gzip.decompress(data)
    if data[:5] not in (b"RDX3\n", b"RDX2\n"):
        raise RFormatError("not an R data file (no RDX2/RDX3 header)")
    r = _Reader(data[5:])
    if r.raw(2) != b"X\n":
        raise RFormatError("only the XDR format is read")
    version = r.int()
    r.int()                                       # writer's R version
    r.int()                                       # minimal reader version
    if version == 3:
        r.raw(r.int())                            # native encoding
    top = r.item()
    if not isinstance(top, _Pairlist):
        raise RFormatError("top level is not a pairlist of saved objects")
    return {k: v for k, v in top.items}


def load(path):
    with open(path, "rb") as f:
        return loads(f.read())


# ------------------------------------------------------------------ writer

class _Writer:
    def __init__(self):
        self.out = io.BytesIO()

    def int(self, v):
        self.out.write(struct.pack(">i", v))

    def flags(self, t, has_attr=False, has_tag=False, levels=0, is_obj=False):
        self.int(t | (IS_OBJECT if is_obj else 0) | (HAS_ATTR if has_attr else 0) | (HAS_TAG if has_tag else 0)
                 | (levels << 12))

    def length(self, n):
        if n < 2 ** 31 - 1:
            self.int(n)
        else:
            self.int(-1)
            self.int(n >> 32)
            self.int(n & 0xFFFFFFFF)

    def charsxp(self, s):
        if s is None:
            self.flags(CHARSXP)
            self.int(-1)
            return
        b = s.encode("utf-8")
        self.flags(CHARSXP, levels=ASCII_MASK if all(c < 128 for c in b) else UTF8_MASK)
        self.int(len(b))
        self.out.write(b)

    def symbol(self, name):
        self.flags(SYMSXP)
        self.charsxp(name)

    def attrs(self, attrs):
        items = list(attrs.items())
        for i, (k, v) in enumerate(items):
            self.flags(LISTSXP, has_tag=True)
            self.symbol(k)
            self.item(v)
        self.int(NILVALUE_SXP)

    def item(self, v):
        attrs = {}
        if isinstance(v, RObj):
            attrs = dict(v.attrs)
            v = v.value
        if v is None:
            if attrs:
                raise ValueError("NULL cannot carry attributes")
            self.int(NILVALUE_SXP)
            return
        if isinstance(v, dict):
            attrs = {"names": list(v.keys()), **attrs}
            v = list(v.values())
        is_chr = isinstance(v, RStr) or (isinstance(v, (list, tuple)) and any(isinstance(x, str) for x in v)
                                         and all(isinstance(x, str) or x is None for x in v))
        if isinstance(v, (list, tuple)) and not is_chr:
            self.flags(VECSXP, has_attr=bool(attrs), is_obj="class" in attrs)
            self.length(len(v))
            for x in v:
                self.item(x)
        elif isinstance(v, (str, list, tuple)) or (isinstance(v, np.ndarray) and v.dtype.kind in "US"):
            seq = [v] if isinstance(v, str) else list(np.asarray(v).ravel(order="F")) if isinstance(v, np.ndarray) \
                else list(v)
            if isinstance(v, np.ndarray) and v.ndim == 2:
                attrs = {"dim": np.array(v.shape, np.int32), **attrs}
            self.flags(STRSXP, has_attr=bool(attrs), is_obj="class" in attrs)
            self.length(len(seq))
            for x in seq:
                self.charsxp(None if x is None else str(x))
        else:
            a = np.asarray(v)
            if a.ndim == 2:
                attrs = {"dim": np.array(a.shape, np.int32), **attrs}
            elif a.ndim > 2:
                raise ValueError("arrays of more than 2 dimensions are not written")
            flat = a.ravel(order="F")
            if a.dtype == np.bool_ or isinstance(v, _Lgl):
                t, data = LGLSXP, np.where(flat.astype(np.int32) < 0, NA_INTEGER, flat.astype(np.int32)).astype(">i4")
            elif a.dtype.kind in "iu":
                t, data = INTSXP, flat.astype(">i4")
            elif a.dtype.kind == "f":
                t, data = REALSXP, flat.astype(">f8")
            else:
                raise ValueError(f"cannot write {a.dtype} to R")
            self.flags(t, has_attr=bool(attrs), is_obj="class" in attrs)
            self.length(flat.size)
            self.out.write(data.tobytes())
        if attrs:
            self.attrs(attrs)


def dumps(objects):
    """The gzip'd "RDX3" body that save() writes for {name: object}."""
    w = _Writer()
    w.out.write(b"RDX3\nX\n")
    w.int(3)
    w.int(R_VERSION)
    w.int(MIN_READER)
    enc = b"UTF-8"
    w.int(len(enc))
    w.out.write(enc)
    for name, obj in objects.items():
        w.flags(LISTSXP, has_tag=True)
        w.symbol(name)
        w.item(obj)
    w.int(NILVALUE_SXP)
    return gzip.compress(w.out.getvalue(), mtime=0)


def save(path, objects):
    with open(path, "wb") as f:
        f.write(dumps(objects))
