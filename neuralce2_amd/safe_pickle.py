"""Data-only reader for the pickle streams inside ``.ce`` and ``model.pkl`` files.

The reference writes these files with ``pickle`` (icet's ``ClusterExpansion.write``, neuralce/nce/mc_model.py:20-21
reads model.pkl with ``pickle.load``).  Unpickling executes code named by the file, so this module never unpickles:
it tokenises the stream with ``pickletools.genops`` (a disassembler) and interprets the opcodes on a small stack of
plain Python values.  References to classes and functions become inert ``Global`` markers; the only ones given a
meaning are numpy's array/dtype/scalar reconstructors, whose raw bytes are turned into arrays with
``numpy.frombuffer``.  Any other object is returned as an ``Obj`` holding its name, constructor arguments and state
dict, with nothing called.
"""
from __future__ import annotations

import pickletools
import numpy as np


class Global:
    __slots__ = ("module", "name")

    def __init__(self, module, name):
        self.module, self.name = module, name

    @property
    def qualname(self):
        return f"{self.module}.{self.name}"

    def __repr__(self):
        return f"Global({self.qualname})"


class Obj:
    """An object of the stream that is not a plain value or a numpy array: name, args, state; nothing executed."""

    def __init__(self, cls, args=()):
        self.cls = cls
        self.args = args
        self.state = None

    def __getitem__(self, k):
        return self.state[k]

    def get(self, k, default=None):
        return self.state.get(k, default) if isinstance(self.state, dict) else default

    def __repr__(self):
        return f"Obj({self.cls.qualname if isinstance(self.cls, Global) else self.cls})"


class _ArrayStub:
    def __init__(self):
        self.value = None


class _DtypeStub:
    def __init__(self, code):
        self.code = code
        self.order = "<"

    def dtype(self):
        dt = np.dtype(self.code)
        if dt.itemsize > 1 and self.order in "<>":
            dt = dt.newbyteorder(self.order)
        return dt


_NUMPY_RECON = {"numpy.core.multiarray._reconstruct", "numpy._core.multiarray._reconstruct"}
_NUMPY_SCALAR = {"numpy.core.multiarray.scalar", "numpy._core.multiarray.scalar"}
_MARK = object()


def _markpos(stack):
    for i in range(len(stack) - 1, -1, -1):
        if stack[i] is _MARK:
            return i
    raise ValueError("pickle MARK not found")


def loads(data: bytes):
    stack = []
    memo = {}
    for op, arg, _pos in pickletools.genops(data):
        name = op.name
        if name in ("PROTO", "FRAME"):
            continue
        if name == "STOP":
            break
        if name == "MARK":
            stack.append(_MARK); continue
        if name in ("MEMOIZE",):
            memo[len(memo)] = stack[-1]; continue
        if name in ("BINPUT", "LONG_BINPUT", "PUT"):
            memo[int(arg)] = stack[-1]; continue
        if name in ("BINGET", "LONG_BINGET", "GET"):
            stack.append(memo[int(arg)]); continue
        if name in ("EMPTY_DICT",):
            stack.append({}); continue
        if name in ("EMPTY_LIST",):
            stack.append([]); continue
        if name in ("EMPTY_TUPLE",):
            stack.append(()); continue
        if name in ("NONE",):
            stack.append(None); continue
        if name in ("NEWTRUE",):
            stack.append(True); continue
        if name in ("NEWFALSE",):
            stack.append(False); continue
        if name in ("BININT", "BININT1", "BININT2", "LONG1", "LONG4", "INT", "LONG", "BINFLOAT", "FLOAT",
                    "SHORT_BINUNICODE", "BINUNICODE", "BINUNICODE8", "UNICODE", "STRING", "BINSTRING",
                    "SHORT_BINSTRING"):
            stack.append(arg); continue
        if name in ("SHORT_BINBYTES", "BINBYTES", "BINBYTES8", "BYTEARRAY8"):
            stack.append(bytes(arg)); continue
        if name in ("TUPLE1", "TUPLE2", "TUPLE3"):
            n = {"TUPLE1": 1, "TUPLE2": 2, "TUPLE3": 3}[name]
            items = tuple(stack[-n:]); del stack[-n:]; stack.append(items); continue
        if name in ("TUPLE", "LIST", "DICT", "FROZENSET"):
            i = _markpos(stack)
            items = stack[i + 1:]; del stack[i:]
            if name == "TUPLE": stack.append(tuple(items))
            elif name == "LIST": stack.append(list(items))
            elif name == "FROZENSET": stack.append(frozenset(items))
            else: stack.append({items[j]: items[j + 1] for j in range(0, len(items), 2)})
            continue
        if name in ("APPEND",):
            v = stack.pop(); stack[-1].append(v); continue
        if name in ("APPENDS", "ADDITEMS"):
            i = _markpos(stack)
            items = stack[i + 1:]; del stack[i:]
            tgt = stack[-1]
            if isinstance(tgt, list): tgt.extend(items)
            else: raise ValueError("ADDITEMS on a non-list")
            continue
        if name in ("EMPTY_SET",):
            stack.append([]); continue
        if name in ("SETITEM",):
            v = stack.pop(); k = stack.pop(); stack[-1][k] = v; continue
        if name in ("SETITEMS",):
            i = _markpos(stack)
            items = stack[i + 1:]; del stack[i:]
            d = stack[-1]
            for j in range(0, len(items), 2):
                d[items[j]] = items[j + 1]
            continue
        if name == "GLOBAL":
            mod, nm = arg.split(" ", 1) if " " in arg else arg.split("\n", 1)
            stack.append(Global(mod, nm)); continue
        if name == "STACK_GLOBAL":
            nm = stack.pop(); mod = stack.pop(); stack.append(Global(mod, nm)); continue
        if name in ("REDUCE", "NEWOBJ", "NEWOBJ_EX"):
            if name == "NEWOBJ_EX":
                stack.pop()
            args = stack.pop(); fn = stack.pop()
            q = fn.qualname if isinstance(fn, Global) else ""
            if q in _NUMPY_RECON:
                stack.append(_ArrayStub())
            elif q == "numpy.dtype":
                stack.append(_DtypeStub(args[0]))
            elif q in _NUMPY_SCALAR:
                dt = args[0].dtype() if isinstance(args[0], _DtypeStub) else np.dtype(args[0])
                stack.append(np.frombuffer(args[1], dtype=dt)[0].item() if len(args) > 1 else None)
            else:
                stack.append(Obj(fn, args))
            continue
        if name == "BUILD":
            state = stack.pop(); obj = stack[-1]
            if isinstance(obj, _ArrayStub):
                _ver, shape, dt, fortran, raw = state
                dtype = dt.dtype() if isinstance(dt, _DtypeStub) else np.dtype(dt)
                if isinstance(raw, (bytes, bytearray)):
                    arr = np.frombuffer(raw, dtype=dtype).copy()
                else:                                    # object arrays: a list of values
                    arr = np.array(raw, dtype=object)
                arr = arr.reshape(shape, order="F" if fortran else "C")
                arr = arr.astype(dtype.newbyteorder("="), copy=False) if dtype.kind in "iufb" else arr
                stack[-1] = arr
                for k, v in list(memo.items()):          # memo entries taken before BUILD point at the stub
                    if v is obj:
                        memo[k] = arr
            elif isinstance(obj, _DtypeStub):
                if isinstance(state, tuple) and len(state) > 1 and isinstance(state[1], str):
                    obj.order = state[1]
            elif isinstance(obj, Obj):
                obj.state = state
            elif isinstance(obj, dict) and isinstance(state, dict):
                obj.update(state)
            else:
                raise ValueError(f"BUILD on {type(obj)}")
            continue
        if name == "POP":
            stack.pop(); continue
        if name == "POP_MARK":
            i = _markpos(stack); del stack[i:]; continue
        if name == "DUP":
            stack.append(stack[-1]); continue
        raise ValueError(f"unsupported pickle opcode {name}")
    if len(stack) != 1:
        raise ValueError(f"malformed pickle stream (stack depth {len(stack)})")
    return stack[0]
