"""Loads a physim transform plugin the way the reference host does.

Mirrors physim-core/src/plugin/discover.rs:327-394 (ABI check,
register_plugin, <el>_register with the host string allocator) and
physim-core/src/plugin/transform.rs:57-151 (TransformElementHandler: load via
<el>_get_api + init(JSON), transform, destroy, property descriptions,
messages). Used to exercise libphysim_gpu.so through the exact symbols a
physim host resolves.
"""
import ctypes
import json

import numpy as np

from ._lib import ACCEL, ENTITY, library_path


class CMessage(ctypes.Structure):
    # physim-core/src/messages.rs:38-45
    _fields_ = [
        ("priority", ctypes.c_int),
        ("topic", ctypes.c_char_p),
        ("message", ctypes.c_char_p),
        ("sender_id", ctypes.c_size_t),
        ("origin", ctypes.c_int),
    ]


ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_char_p)
INIT_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t)
TRANSFORM_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                ctypes.c_void_p, ctypes.c_size_t)
DESTROY_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p)
PROPS_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ALLOC_FN)
RECV_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(CMessage))
POST_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p)


class TransformElementAPI(ctypes.Structure):
    # physim-core/src/plugin/transform.rs:32-50
    _fields_ = [
        ("init", INIT_FN),
        ("transform", TRANSFORM_FN),
        ("destroy", DESTROY_FN),
        ("get_property_descriptions", PROPS_FN),
        ("recv_message", RECV_FN),
        ("post_configuration_messages", POST_FN),
    ]


class ElementMetaFFI(ctypes.Structure):
    # physim-core/src/plugin/meta.rs:246-257
    _fields_ = [("kind", ctypes.c_int)] + [
        (f, ctypes.c_void_p) for f in ("name", "plugin", "version", "license", "author", "blurb",
                                       "repo")
    ]


_libc = ctypes.CDLL(None)
_libc.malloc.restype = ctypes.c_void_p
_libc.malloc.argtypes = [ctypes.c_size_t]
_libc.free.argtypes = [ctypes.c_void_p]


@ALLOC_FN
def host_alloc_string(s):
    """The host allocator plugins must use for strings they hand back
    (physim-core/src/plugin/meta.rs:288-306)."""
    if s is None:
        return None
    b = bytes(s).replace(b"\0", b"\xff") + b"\0"
    p = _libc.malloc(len(b))
    ctypes.memmove(p, b, len(b))
    return p


def _take_string(p):
    if not p:
        return None
    s = ctypes.string_at(p).decode("utf-8", "replace")
    _libc.free(p)
    return s


class PluginLibrary:
    """One plugin shared library, scanned as discover.rs scans it."""

    def __init__(self, path=None):
        self.path = path or library_path()
        self.lib = ctypes.CDLL(self.path)
        self.lib.get_plugin_abi_info.restype = ctypes.c_char_p
        self.lib.register_plugin.restype = ctypes.c_char_p
        self.lib.set_callback_target.argtypes = [ctypes.c_void_p]

    def abi_info(self):
        return self.lib.get_plugin_abi_info().decode()

    def element_names(self):
        return self.lib.register_plugin().decode().split(",")

    def meta(self, name):
        f = getattr(self.lib, f"{name}_register")
        f.restype = ElementMetaFFI
        f.argtypes = [ALLOC_FN]
        m = f(host_alloc_string)
        out = {"kind": m.kind}
        for field in ("name", "plugin", "version", "license", "author", "blurb", "repo"):
            out[field] = _take_string(getattr(m, field))
        return out

    def set_callback_target(self, target):
        self.lib.set_callback_target(target)

    def load(self, name, properties=None):
        return TransformElementHandler(self, name, properties or {})


class TransformElementHandler:
    """physim-core/src/plugin/transform.rs:52-151."""

    def __init__(self, plugin, name, properties):
        get_api = getattr(plugin.lib, f"{name}_get_api")
        get_api.restype = ctypes.POINTER(TransformElementAPI)
        self.api = get_api().contents
        self.name = name
        cfg = json.dumps(properties).encode()
        self._cfg = ctypes.create_string_buffer(cfg, len(cfg))
        inst = self.api.init(ctypes.cast(self._cfg, ctypes.c_void_p), len(cfg))
        if not inst:
            raise ValueError(f"NullElement: {name} rejected {properties!r}")
        self.instance = inst

    @classmethod
    def from_raw_json(cls, plugin, name, raw):
        """init() with raw bytes (for malformed-JSON checks); returns the
        instance pointer or None."""
        get_api = getattr(plugin.lib, f"{name}_get_api")
        get_api.restype = ctypes.POINTER(TransformElementAPI)
        api = get_api().contents
        buf = ctypes.create_string_buffer(raw, len(raw))
        inst = api.init(ctypes.cast(buf, ctypes.c_void_p), len(raw))
        if inst:
            api.destroy(inst)
        return inst

    def transform(self, state, acceleration):
        """acceleration += this element's contribution (both host AoS)."""
        assert state.dtype == ENTITY and acceleration.dtype == ACCEL
        assert state.flags["C_CONTIGUOUS"] and acceleration.flags["C_CONTIGUOUS"]
        n = len(state)
        self.api.transform(self.instance, state.ctypes.data, n, acceleration.ctypes.data, n)

    def get_property_descriptions(self):
        p = self.api.get_property_descriptions(self.instance, host_alloc_string)
        s = _take_string(p)
        if s is None:
            raise RuntimeError("Unable to load descriptions of properties")
        return json.loads(s)

    def recv_message(self, topic, message, priority=2, sender=0):
        m = CMessage(priority, topic.encode(), message.encode(), sender, 0)
        self.api.recv_message(self.instance, ctypes.byref(m))

    def post_configuration_messages(self):
        self.api.post_configuration_messages(self.instance)

    def close(self):
        if getattr(self, "instance", None):
            self.api.destroy(self.instance)
            self.instance = None

    def __del__(self):
        self.close()


def zero_acc(n):
    return np.zeros(n, ACCEL)
