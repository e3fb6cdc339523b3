"""Raw scene descriptor validation and decoding (CPU: no device needed).

The library's decoder follows native/docs/scene-format.md and the records of
native/include/mutracer-scene.h; `mtx_descriptor_validate` runs it without a
context. Each case breaks one rule the format states and expects the status
the reference's validator returns for it (INVALID_ARGUMENT), or UNAVAILABLE
for a well-formed feature this renderer does not draw.
"""
import numpy as np
import pytest

from mutracer_amd import abi, context, descriptor, scene


@pytest.fixture(scope="module")
def cornell():
    return scene.cornell_box()


def _words(data):
    return np.frombuffer(data, np.uint32).copy()


def _status(words):
    try:
        context.validate(words.tobytes())
        return abi.MTX_OK
    except abi.MtxError as e:
        return e.code


def test_encode_summary(cornell):
    d = descriptor.encode(cornell, 64, 48)
    assert context.validate(d) == dict(width=64, height=48, blas_meshes=7, instances=7, materials=3, lights=1,
                                       env_width=0, env_height=0)
    h = descriptor.header(d)
    assert h[0] == 8 and h[1] * 4 == len(d)
    field = scene.instanced_field(n_instances=4, freq=4, env_size=(32, 16))
    s = context.validate(descriptor.encode(field, 20, 10))
    assert s["instances"] == 5 and s["env_width"] == 32 and s["env_height"] == 16 and s["materials"] == len(
        field.materials)


@pytest.mark.parametrize("word,value", [(0, 4), (0, 9), (2, 0), (3, 9000), (4, 0), (5, 0), (15, 1), (30, 7)])
def test_header_rules(cornell, word, value):
    w = _words(descriptor.encode(cornell, 32, 32))
    w[word] = value
    assert _status(w) == abi.MTX_INVALID_ARGUMENT


def test_length_and_offsets(cornell):
    w = _words(descriptor.encode(cornell, 32, 32))
    bad = w.copy()
    bad[1] += 1  # total_words disagrees with the buffer
    assert _status(bad) == abi.MTX_INVALID_ARGUMENT
    for table in (9, 10, 11, 13, 14):
        bad = w.copy()
        bad[table] = 3  # inside the header
        assert _status(bad) == abi.MTX_INVALID_ARGUMENT
        bad[table] = w.size  # past the end
        assert _status(bad) == abi.MTX_INVALID_ARGUMENT
    with pytest.raises(abi.MtxError):
        context.validate(descriptor.encode(cornell, 32, 32)[:-2])


def test_record_rules(cornell):
    w = _words(descriptor.encode(cornell, 32, 32))
    geo, mesh, mat, light, cam = (int(w[i]) for i in (9, 10, 11, 12, 14))
    f = w.view(np.float32)
    cases = []
    c = w.copy(); c[int(w[geo + 3])] = 10 ** 6; cases.append(("index out of range", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c[geo + 1] = 5; cases.append(("index count % 3", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c[geo + 10] = 1; cases.append(("geometry reserved", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c.view(np.float32)[int(w[geo + 2])] = np.inf; cases.append(("non-finite position", c,
                                                                              abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c[mesh] = 99; cases.append(("mesh geometry", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c[mesh + 1] = 99; cases.append(("mesh material", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c[mesh + 20] = 2; cases.append(("mesh flags", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c.view(np.float32)[mesh + 4 + 15] = 2.0; cases.append(("non-affine mesh", c,
                                                                         abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c.view(np.float32)[mesh + 4:mesh + 7] = 0.0; cases.append(("singular mesh", c,
                                                                             abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c[mat + 14] = 2; cases.append(("thin_walled 2", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c[mat + 22] = 1; cases.append(("material reserved", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c[mat + 15] = 5; cases.append(("texture index", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c[mat + 21] = 40; cases.append(("material graph", c, abi.MTX_UNAVAILABLE))
    c = w.copy(); c.view(np.float32)[mat + 13] = 0.5; cases.append(("opacity", c, abi.MTX_UNAVAILABLE))
    c = w.copy(); c[light] = 6; cases.append(("light type", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c.view(np.float32)[light + 5] = 0.0; cases.append(("rect width", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c[light + 8] = 1; cases.append(("rect texture", c, abi.MTX_UNAVAILABLE))
    c = w.copy(); c[light + 11] = 3; cases.append(("light reserved", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c[cam] = 2; cases.append(("projection", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c.view(np.float32)[cam + 1] = 0.0; cases.append(("fov", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c[cam + 13] = 1; cases.append(("camera reserved", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c[22] = 1; cases.append(("bevel", c, abi.MTX_UNAVAILABLE))
    c = w.copy(); c[22] = 8; cases.append(("unknown quality flag", c, abi.MTX_INVALID_ARGUMENT))
    c = w.copy(); c.view(np.float32)[16] = 0.5; cases.append(("background color", c, abi.MTX_UNAVAILABLE))
    assert f[mat + 13] == 1.0
    for name, words, code in cases:
        assert _status(words) == code, name


def _extended(cornell, ext_words):
    """The Cornell descriptor with one extension entry added to material 0."""
    w = _words(descriptor.encode(cornell, 16, 16))
    mat = int(w[11])
    w[mat + 19] = 1
    w[mat + 20] = w.size
    w = np.concatenate([w, np.asarray(ext_words, np.uint32)])
    w[1] = w.size
    return w


def test_material_extensions(cornell):
    one = np.float32(1.0).view(np.uint32)
    assert _status(_extended(cornell, [25, np.float32(0.3).view(np.uint32), 0, 0])) == abi.MTX_OK  # coat weight
    assert _status(_extended(cornell, [38, 1, 0, 0])) == abi.MTX_OK                               # double sided
    assert _status(_extended(cornell, [77, 0, 0, 0])) == abi.MTX_INVALID_ARGUMENT                 # unknown id
    assert _status(_extended(cornell, [38, 2, 0, 0])) == abi.MTX_INVALID_ARGUMENT                 # bool 2
    assert _status(_extended(cornell, [25, one, 1, 0])) == abi.MTX_INVALID_ARGUMENT               # unused word
    assert _status(_extended(cornell, [31, one, 0, 0])) == abi.MTX_UNAVAILABLE                    # thin film
    assert _status(_extended(cornell, [39, 0, 0, 0])) == abi.MTX_INVALID_ARGUMENT                 # no texture 0
    assert _status(_extended(cornell, [39, 0xFFFFFFFF, 0, 0])) == abi.MTX_OK                      # cleared slot
    w = _extended(cornell, [25, one, 0, 0])
    mat = int(w[11])
    w[mat + 19] = 2
    w = np.concatenate([w, np.asarray([25, one, 0, 0], np.uint32)])
    w[1] = w.size
    assert _status(w) == abi.MTX_INVALID_ARGUMENT  # the same id twice


def test_hidden_meshes_batches_and_material_tables(cornell):
    w = _words(descriptor.encode(cornell, 16, 16))
    mesh = int(w[10])
    hidden = w.copy()
    hidden[mesh + 20] = 1
    assert context.validate(hidden.tobytes())["instances"] == 6
    # Batches: mesh 0 expanded to three local matrices, the others ordinary.
    n_mesh = int(w[5])
    eye = np.eye(4, dtype=np.float32).reshape(16)
    mats = np.concatenate([eye, eye, eye]).view(np.uint32)
    w2 = np.concatenate([w, mats])
    batch = np.zeros((n_mesh, 4), np.uint32)
    batch[:, 0] = 0xFFFFFFFF
    batch[0, 0], batch[0, 1] = 3, w.size
    w2 = np.concatenate([w2, batch.reshape(-1)])
    w2[23] = w.size + mats.size
    w2[1] = w2.size
    s = context.validate(w2.tobytes())
    assert s["instances"] == 9 and s["blas_meshes"] == 7
    empty = w2.copy()
    empty[w.size + mats.size] = 0
    empty[w.size + mats.size + 1] = 0
    assert context.validate(empty.tobytes())["instances"] == 6


def _two_triangle_scene():
    s = scene.Scene()
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [2, 0, 0], [3, 0, 0], [2, 1, 0]], np.float32)
    s.add_instance(s.add_mesh(pos, np.arange(6, dtype=np.uint32)), np.eye(4).reshape(16), 0)
    s.materials = [scene.Material(), scene.Material(base_color=(0.5, 0.1, 0.1))]
    s.camera = scene.Camera(eye=(1.5, 0.5, 4.0), target=(1.5, 0.5, 0.0))
    return s


def _with_material_slots(w, slots, table):
    geo, mesh = int(w[9]), int(w[10])
    out = np.concatenate([w, np.asarray(slots, np.uint32), np.asarray(table, np.uint32)])
    out[geo + 9] = w.size
    out[mesh + 2], out[mesh + 3] = w.size + len(slots), len(table)
    out[1] = out.size
    return out


def test_multi_material_geometry_splits():
    """Per-vertex material indices select entries of the mesh's material
    table; the decoder builds one BLAS per material of the geometry."""
    w = _words(descriptor.encode(_two_triangle_scene(), 16, 16))
    s = context.validate(_with_material_slots(w, [0, 0, 0, 1, 1, 1], [0, 1]).tobytes())
    assert s["blas_meshes"] == 2 and s["instances"] == 2
    s = context.validate(_with_material_slots(w, [1, 1, 1, 1, 1, 1], [0, 1]).tobytes())
    assert s["blas_meshes"] == 1 and s["instances"] == 1
    # A triangle whose vertices select different materials is rejected, and so
    # is a table entry past the material table.
    assert _status(_with_material_slots(w, [0, 1, 0, 1, 1, 1], [0, 1])) == abi.MTX_INVALID_ARGUMENT
    assert _status(_with_material_slots(w, [0, 0, 0, 1, 1, 1], [0, 7])) == abi.MTX_INVALID_ARGUMENT


def test_version5_light_records(cornell):
    """Version 5 packs 28-word light records (no radius/length/normalize)."""
    w = _words(descriptor.encode(cornell, 16, 16))
    light = int(w[12])
    v5 = np.concatenate([w[:light], w[light:light + 28], w[light + 32:]])
    v5[0] = 5
    for i in (9, 10, 11, 13, 14):  # tables after the light record move up by 4 words
        if v5[i] > light:
            v5[i] -= 4
    # The geometry words (positions/indices/normals) and extension offsets
    # stored inside records also point past the light table only if laid out
    # after it; the encoder writes arrays first, so only table offsets move.
    v5[1] = v5.size
    assert context.validate(v5.tobytes())["lights"] == 1


def test_material_textures_and_alpha_resolution():
    """Material textures travel through the descriptor (basic-record slots,
    extension slots, UV sets, texture records); alpha resolves as
    resolveMaterialAlphaMode (MaterialState.ts:251-264): mask materials
    (with or without coverage maps) are accepted, blend coverage is refused
    as UNAVAILABLE; texture features outside the row (UDIM sets, coat normal
    maps) are UNAVAILABLE too."""
    room = scene.textured_room()
    s = context.validate(descriptor.encode(room, 24, 16))
    assert s["materials"] == len(room.materials) and s["instances"] == len(room.mesh_index)

    def status_of(edit):
        sc = scene.textured_room()
        edit(sc)
        try:
            context.validate(descriptor.encode(sc, 8, 8))
            return abi.MTX_OK
        except abi.MtxError as e:
            return e.code

    def opacity(sc):
        sc.materials[0].opacity = 0.5
    assert status_of(opacity) == abi.MTX_UNAVAILABLE

    def opaque_override(sc):
        sc.materials[0].opacity = 0.5
        sc.materials[0].alpha_mode = "opaque"
    assert status_of(opaque_override) == abi.MTX_OK

    def mask(sc):
        sc.materials[0].alpha_cutoff = 0.5
    assert status_of(mask) == abi.MTX_OK

    def mask_with_maps(sc):
        sc.materials[0].alpha_cutoff = 0.5
        sc.materials[0].alpha_map = 1
        sc.materials[0].map_alpha_usage = "alpha"
        sc.materials[0].opacity = 0.75
    assert status_of(mask_with_maps) == abi.MTX_OK

    def blend(sc):
        sc.materials[0].alpha_mode = "blend"
    assert status_of(blend) == abi.MTX_UNAVAILABLE

    def alpha_map(sc):
        sc.materials[0].alpha_map = 1
    assert status_of(alpha_map) == abi.MTX_UNAVAILABLE

    def map_alpha(sc):
        sc.materials[0].map_alpha_usage = "alpha"
    assert status_of(map_alpha) == abi.MTX_UNAVAILABLE
    w = _words(descriptor.encode(room, 8, 8))
    tex = int(w[13])
    udim = w.copy()
    udim[tex + 12] = 1  # texture 0 (the floor's base map) as a UDIM set
    assert _status(udim) == abi.MTX_UNAVAILABLE
    bad = w.copy()
    bad[tex + 10] = 3  # wrap mode out of range
    assert _status(bad) == abi.MTX_INVALID_ARGUMENT
    assert _status(_extended(scene.cornell_box(), [74, 0xFFFFFFFF, 0, 0])) == abi.MTX_OK
    room_words = _words(descriptor.encode(room, 8, 8))
    mat = int(room_words[11])
    room_words[mat + 16] = 40  # normal map slot naming a texture past the table
    assert _status(room_words) == abi.MTX_INVALID_ARGUMENT
