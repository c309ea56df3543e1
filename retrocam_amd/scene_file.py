"""Scene files and the editable scene: SceneSerializer
(src/main/java/com/retrocam/scene/SceneSerializer.java:24-217), SceneEditor
(SceneEditor.java:16-134) and SceneObject (SceneObject.java:10-98).

`load` reads a file with the Java reader's rules (librc_host's rcs_json_parse),
so a file gives the materials and objects the Java host reads from it; `save`
writes the text SceneSerializer.save writes (same layout, floats printed as
Java's Float.toString prints them). `SceneEditor.build` is buildScene: the
materials, then each object's box or sphere (rotated about its position when
an angle is set), then the BVH and light list.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np

from . import _native as N
from .scene import Material, SceneArrays, SceneBuilder

f32 = np.float32


class SceneFileError(ValueError):
    pass


# ---- Float.toString ------------------------------------------------------------------

def _round_to_f32(q: Fraction) -> float:
    """The float32 nearest to the rational q (ties to even), without double rounding."""
    return float(f32(float(q))) if q == 0 else _nearest_f32(q)


def _nearest_f32(q: Fraction) -> float:
    a = f32(float(q))  # within one float32 ulp of the answer
    best = None
    for c in (np.nextafter(a, f32(-np.inf)), a, np.nextafter(a, f32(np.inf))):
        if not np.isfinite(c):
            continue
        d = abs(Fraction(float(c)) - q)
        key = (d, int(np.array(c, f32).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, c)
    return float(best[1])


def _digits_exp(x: float):
    """Shortest decimal digits and exponent that read back as x (float32), per
    Float.toString (JDK 19+): the shortest decimals that round to x; when one
    digit suffices, the closest among the one- and two-digit ones."""
    s = np.format_float_scientific(f32(x), unique=True, trim="-")
    mant, e = s.split("e")
    digits = mant.replace(".", "").lstrip("-")
    e = int(e)
    if len(digits) > 1:
        return digits, e
    xq = Fraction(float(f32(x)))
    best = (abs(Fraction(int(digits)) * Fraction(10) ** e - xq), digits, e)
    for ee in (e, e - 1):
        for m in range(10, 100):
            q = Fraction(m) * Fraction(10) ** (ee - 1)
            if _round_to_f32(q) == float(f32(x)):
                d = abs(q - xq)
                cand = (d, str(m).rstrip("0") or "0", ee)
                if d < best[0] or (d == best[0] and len(cand[1]) < len(best[1])):
                    best = cand
    return best[1], best[2]


def java_float_str(x) -> str:
    """Float.toString(float) (Java 19+ specification)."""
    x = f32(x)
    if np.isnan(x):
        return "NaN"
    if np.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    if x == 0:
        return "-0.0" if np.signbit(x) else "0.0"
    sign = "-" if x < 0 else ""
    digits, e = _digits_exp(abs(float(x)))
    if 1e-3 <= abs(float(x)) < 1e7:
        if e >= 0:
            ip = (digits + "0" * (e + 1))[: e + 1] if len(digits) <= e + 1 else digits[: e + 1]
            fp = digits[e + 1:] or "0"
            return f"{sign}{ip}.{fp}"
        return f"{sign}0.{'0' * (-e - 1)}{digits}"
    return f"{sign}{digits[0]}.{digits[1:] or '0'}E{e}"


def _jstr(s: str) -> str:  # SceneSerializer.jStr, :215-217
    return '"' + s.replace("\\", "\\\\").replace('"', '\\"') + '"'


# ---- SceneObject / SceneEditor ---------------------------------------------------------

@dataclass
class SceneObject:
    """SceneObject (SceneObject.java:10-98)."""
    name: str
    type: str = "BOX"  # "BOX" or "SPHERE"
    px: float = 0.0
    py: float = 0.0
    pz: float = 0.0
    sx: float = 1.0
    sy: float = 1.0
    sz: float = 1.0
    rx: float = 0.0
    ry: float = 0.0
    rz: float = 0.0
    stacks: int = 24
    slices: int = 24
    materialIndex: int = 0

    _PROPS = {"pos.x": "px", "pos.y": "py", "pos.z": "pz", "scl.x": "sx", "scl.y": "sy", "scl.z": "sz",
              "rot.x": "rx", "rot.y": "ry", "rot.z": "rz"}

    # Keyframeable (:45-78): no clamps; the caller re-builds the scene after changes
    def keyframeable_property_names(self):
        return list(self._PROPS)

    def get_keyframeable_property(self, name: str) -> float:
        return float(f32(getattr(self, self._PROPS[name]))) if name in self._PROPS else 0.0

    def set_keyframeable_property(self, name: str, value: float):
        if name in self._PROPS:
            setattr(self, self._PROPS[name], float(f32(value)))

    def to_c(self) -> N.RcsJsonObject:
        o = N.RcsJsonObject()
        o.type = N.RCS_OBJ_SPHERE if self.type.upper() == "SPHERE" else N.RCS_OBJ_BOX
        o.pos[:] = [float(self.px), float(self.py), float(self.pz)]
        o.scale[:] = [float(self.sx), float(self.sy), float(self.sz)]
        o.rot[:] = [float(self.rx), float(self.ry), float(self.rz)]
        o.stacks, o.slices, o.materialIndex = int(self.stacks), int(self.slices), int(self.materialIndex)
        return o

    @staticmethod
    def default_box(mat: int) -> "SceneObject":  # :92-94
        return SceneObject("New Box", "BOX", 0.0, 1.0, 0.0, 1.0, 1.0, 1.0, materialIndex=mat)

    @staticmethod
    def default_sphere(mat: int) -> "SceneObject":  # :96-98
        return SceneObject("New Sphere", "SPHERE", 0.0, 1.0, 0.0, 0.5, 0.5, 0.5, materialIndex=mat)


@dataclass
class SceneEditor:
    """SceneEditor (SceneEditor.java:16-134): named materials and objects with a
    dirty flag; build() compiles them into the arrays rc_upload_scene takes."""
    materials: list = field(default_factory=list)
    mat_names: list = field(default_factory=list)
    objects: list = field(default_factory=list)
    dirty: bool = True

    def mark_dirty(self):
        self.dirty = True

    def clear_dirty(self):
        self.dirty = False

    def add_material(self, name: str, m: Material) -> int:
        self.materials.append(m)
        self.mat_names.append(name)
        self.dirty = True
        return len(self.materials) - 1

    def remove_material(self, idx: int):  # :57-66, remapping object references
        if not 0 <= idx < len(self.materials):
            return
        del self.materials[idx]
        del self.mat_names[idx]
        for o in self.objects:
            if o.materialIndex >= idx:
                o.materialIndex = max(0, o.materialIndex - 1)
        self.dirty = True

    def add_object(self, o: SceneObject):
        self.objects.append(o)
        self.dirty = True

    def remove_object(self, idx: int):
        if 0 <= idx < len(self.objects):
            del self.objects[idx]
            self.dirty = True

    def clear_all(self):
        self.objects.clear()
        self.materials.clear()
        self.mat_names.clear()
        self.dirty = True

    def build(self) -> SceneArrays:
        """buildScene (:101-107) + Scene.build: the BVH and lights."""
        b = SceneBuilder()
        for m in self.materials:
            b.add_material(m)
        lib = N.host_lib()
        for o in self.objects:
            c = o.to_c()
            if lib.rcs_add_object(b._s, C.byref(c)) < 0:
                raise SceneFileError(f"object {o.name!r}: bad parameters")
        return b.build()

    @staticmethod
    def create_default() -> "SceneEditor":  # :111-134, the default Cornell box
        e = SceneEditor()
        white = e.add_material("White", Material.diffuse(0.73, 0.73, 0.73))
        red = e.add_material("Red", Material.diffuse(0.65, 0.05, 0.05))
        green = e.add_material("Green", Material.diffuse(0.12, 0.45, 0.15))
        light = e.add_material("Light", Material.emissive(1.0, 0.95, 0.85, 15.0))
        metal = e.add_material("Metal", Material.metal(0.95, 0.85, 0.55, 0.1))
        glass = e.add_material("Glass", Material.glass(1.5))
        H = 5.0
        two_h = float(f32(2.0) * f32(H))
        for args in (("Floor", "BOX", 0.0, 0.0, 0.0, H, 0.05, H, white),
                     ("Ceiling", "BOX", 0.0, two_h, 0.0, H, 0.05, H, white),
                     ("Left Wall", "BOX", -H, H, 0.0, 0.05, H, H, red),
                     ("Right Wall", "BOX", H, H, 0.0, 0.05, H, H, green),
                     ("Back Wall", "BOX", 0.0, H, -H, H, H, 0.05, white),
                     ("Area Light", "BOX", 0.0, float(f32(two_h) - f32(0.1)), 0.0, 1.5, 0.05, 1.5, light),
                     ("Tall Block", "BOX", -1.5, 1.5, -1.0, 1.0, 1.5, 1.0, white),
                     ("Metal Block", "BOX", 1.5, 1.0, 0.5, 1.0, 1.0, 1.0, metal),
                     ("Glass Sphere", "SPHERE", 0.0, 1.2, 1.5, 1.2, 0.0, 0.0, glass)):
            name, typ, px, py, pz, sx, sy, sz, mat = args
            e.add_object(SceneObject(name, typ, px, py, pz, sx, sy, sz, materialIndex=mat))
        e.dirty = True
        return e


# ---- save / load ---------------------------------------------------------------------

def dumps(editor: SceneEditor) -> str:
    """SceneSerializer.save's text (:24-66)."""
    F = java_float_str
    out = ["{\n", '  "materials": [\n']
    n = len(editor.materials)
    for i, (m, name) in enumerate(zip(editor.materials, editor.mat_names)):
        a, em = m.albedo, m.emission
        out.append("    {\n")
        out.append(f'      "name": {_jstr(name)},\n')
        out.append(f'      "albedo": [{F(a[0])}, {F(a[1])}, {F(a[2])}],\n')
        out.append(f'      "metallic": {F(m.metallic)},\n')
        out.append(f'      "emission": [{F(em[0])}, {F(em[1])}, {F(em[2])}],\n')
        out.append(f'      "emissionStrength": {F(m.emissionStrength)},\n')
        out.append(f'      "roughness": {F(m.roughness)},\n')
        out.append(f'      "ior": {F(m.ior)},\n')
        out.append(f'      "transmission": {F(m.transmission)}\n')
        out.append("    }" + ("," if i < n - 1 else "") + "\n")
    out.append("  ],\n")
    out.append('  "objects": [\n')
    n = len(editor.objects)
    for i, o in enumerate(editor.objects):
        out.append("    {\n")
        out.append(f'      "name": {_jstr(o.name)},\n')
        out.append(f'      "type": {_jstr("SPHERE" if o.type.upper() == "SPHERE" else "BOX")},\n')
        out.append(f'      "px": {F(o.px)}, "py": {F(o.py)}, "pz": {F(o.pz)},\n')
        out.append(f'      "sx": {F(o.sx)}, "sy": {F(o.sy)}, "sz": {F(o.sz)},\n')
        out.append(f'      "rx": {F(o.rx)}, "ry": {F(o.ry)}, "rz": {F(o.rz)},\n')
        out.append(f'      "stacks": {int(o.stacks)}, "slices": {int(o.slices)},\n')
        out.append(f'      "materialIndex": {int(o.materialIndex)}\n')
        out.append("    }" + ("," if i < n - 1 else "") + "\n")
    out.append("  ]\n")
    out.append("}\n")
    return "".join(out)


def save(editor: SceneEditor, path: str):
    with open(path, "w", encoding="utf-8", newline="") as f:
        f.write(dumps(editor))


def loads(text: str | bytes) -> SceneEditor:
    """SceneSerializer.load's result for this text (:72-124)."""
    data = text.encode("utf-8") if isinstance(text, str) else bytes(text)
    lib = N.host_lib()
    st = C.c_int32(0)
    d = lib.rcs_json_parse(data, len(data), C.byref(st))
    if not d:
        raise SceneFileError({N.RCS_JSON_ERR_STRING: "a string field's scan runs past the text",
                              N.RCS_JSON_ERR_OOM: "out of memory"}.get(st.value, f"parse failed ({st.value})"))
    try:
        e = SceneEditor()
        w = np.zeros(16, f32)
        for i in range(lib.rcs_json_material_count(d)):
            n = lib.rcs_json_get_material(d, i, N.fptr(w), None, 0)
            buf = C.create_string_buffer(n + 1)
            lib.rcs_json_get_material(d, i, N.fptr(w), buf, n + 1)
            m = Material(albedo=tuple(float(v) for v in w[0:3]), metallic=float(w[3]),
                         emission=tuple(float(v) for v in w[4:7]), emissionStrength=float(w[7]),
                         roughness=float(w[8]), ior=float(w[9]), transmission=float(w[10]))
            e.add_material(buf.raw[:n].decode("utf-8", errors="surrogateescape"), m)
        o = N.RcsJsonObject()
        for i in range(lib.rcs_json_object_count(d)):
            n = lib.rcs_json_get_object(d, i, C.byref(o), None, 0)
            buf = C.create_string_buffer(n + 1)
            lib.rcs_json_get_object(d, i, C.byref(o), buf, n + 1)
            e.add_object(SceneObject(buf.raw[:n].decode("utf-8", errors="surrogateescape"),
                                     "SPHERE" if o.type == N.RCS_OBJ_SPHERE else "BOX",
                                     *[float(v) for v in o.pos], *[float(v) for v in o.scale],
                                     *[float(v) for v in o.rot], int(o.stacks), int(o.slices),
                                     int(o.materialIndex)))
        e.dirty = True
        return e
    finally:
        lib.rcs_json_free(d)


def load(path: str) -> SceneEditor:
    with open(path, "rb") as f:
        return loads(f.read())


def build_file(path: str) -> SceneArrays:
    """A scene file straight to the arrays rc_upload_scene takes (rcs_json_build)."""
    with open(path, "rb") as f:
        data = f.read()
    lib = N.host_lib()
    st = C.c_int32(0)
    d = lib.rcs_json_parse(data, len(data), C.byref(st))
    if not d:
        raise SceneFileError(f"parse failed ({st.value})")
    try:
        b = SceneBuilder()
        if lib.rcs_json_build(d, b._s) < 0:
            raise SceneFileError("bad object parameters")
        return b.build()
    finally:
        lib.rcs_json_free(d)
