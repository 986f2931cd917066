// athx_ply_parse: the header of a binary PLY file, read into the vertex
// element's layout the ingest kernels take (host only, no device touched).
//
// The grammar is PLY's own (Turk's 1994 format note): a first line that
// starts with "ply", then one keyword line at a time -- "format <name>
// <version>", "comment ...", "obj_info ...", "element <name> <count>",
// "property <type> <name>", "property list <count type> <item type> <name>"
// -- up to "end_header"; the binary data starts right after that line's
// newline. Lines may end in "\r\n". Elements' records follow one another in
// header order, each element's records back to back.
//
// What the reference makes of a header, and which of its refusals a caller
// sees with which ErrorCode and message, is its interface and is kept:
// io::parsePlyHeader (modules/io/src/PlyHeader.cpp:59-128) and the checks of
// io::readSplatPly (modules/io/src/SplatReaders.cpp:48-195), in that order.
// A property of a type the format does not name counts 4 bytes, as there.
#include <charconv>
#include <cstring>
#include <string>
#include <string_view>
#include <vector>

#include "athx.h"

namespace {

// The scalar types of the format (its original names and the sized ones
// later writers use), each with its bytes and the ingest's type code.
struct ScalarType {
    uint32_t bytes;
    int code;  // athx_ply_type, or -1 for a name the format does not have
};

ScalarType scalar_type(std::string_view name) {
    const auto is = [&](const char* a, const char* b) { return name == a || name == b; };
    if (is("float", "float32")) return {4, ATHX_PLY_F32};
    if (is("double", "float64")) return {8, ATHX_PLY_F64};
    if (is("uchar", "uint8")) return {1, ATHX_PLY_U8};
    if (is("char", "int8")) return {1, ATHX_PLY_I8};
    if (is("ushort", "uint16")) return {2, ATHX_PLY_U16};
    if (is("short", "int16")) return {2, ATHX_PLY_I16};
    if (is("uint", "uint32")) return {4, ATHX_PLY_U32};
    if (is("int", "int32")) return {4, ATHX_PLY_I32};
    return {4, -1};
}

// Whitespace-separated words of one header line.
std::vector<std::string_view> words_of(std::string_view line) {
    std::vector<std::string_view> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && (line[i] == ' ' || line[i] == '\t' || line[i] == '\v' || line[i] == '\f')) {
            ++i;
        }
        const size_t start = i;
        while (i < line.size() && !(line[i] == ' ' || line[i] == '\t' || line[i] == '\v' || line[i] == '\f')) {
            ++i;
        }
        if (i > start) {
            out.push_back(line.substr(start, i - start));
        }
    }
    return out;
}

// An element count: its leading decimal digits (none: 0).
uint64_t count_of(std::string_view word) {
    uint64_t v = 0;
    const auto r = std::from_chars(word.data(), word.data() + word.size(), v);
    return r.ec == std::errc() ? v : 0;
}

struct Property {
    std::string name, type;
    bool list = false;
    uint32_t offset = 0;
    ScalarType scalar{};
};

struct Element {
    std::string name;
    uint64_t count = 0;
    uint64_t record_bytes = 0;  // a record's bytes when it holds no list
    bool holds_list = false;
    std::vector<Property> props;

    int index_of(std::string_view prop) const {
        for (size_t k = 0; k < props.size(); ++k) {
            if (props[k].name == prop) {
                return (int)k;
            }
        }
        return -1;
    }
};

struct Header {
    std::string format;
    std::vector<Element> elements;
    uint64_t data_start = 0;  // the first byte after end_header's line
};

// A failure: the status and the reference's reason.
struct Refusal {
    athx_status status = ATHX_OK;
    std::string reason;
};

// Reads the header lines of `text`; the first refusal ends it.
Refusal read_header(std::string_view text, Header& h) {
    size_t pos = 0;
    bool first = true;
    while (pos < text.size()) {
        const size_t nl = text.find('\n', pos);
        const size_t stop = nl == std::string_view::npos ? text.size() : nl;
        std::string_view line = text.substr(pos, stop - pos);
        pos = nl == std::string_view::npos ? text.size() : nl + 1;
        if (!line.empty() && line.back() == '\r') {
            line.remove_suffix(1);
        }
        if (first) {
            if (line.substr(0, 3) != "ply") {
                return {ATHX_INVALID_ARGUMENT, "not a PLY file"};
            }
            first = false;
            continue;
        }
        const std::vector<std::string_view> w = words_of(line);
        if (w.empty()) {
            continue;
        }
        const std::string_view key = w[0];
        if (key == "end_header") {
            h.data_start = pos;
            if (h.format.empty()) {
                return {ATHX_INVALID_ARGUMENT, "PLY header has no format line"};
            }
            return {};
        }
        if (key == "format") {
            h.format = w.size() > 1 ? std::string(w[1]) : std::string();
        } else if (key == "element") {
            Element e;
            e.name = w.size() > 1 ? std::string(w[1]) : std::string();
            e.count = w.size() > 2 ? count_of(w[2]) : 0;
            h.elements.push_back(std::move(e));
        } else if (key == "property") {
            if (h.elements.empty()) {
                return {ATHX_INVALID_ARGUMENT, "a PLY property before any element"};
            }
            Element& e = h.elements.back();
            Property p;
            if (w.size() > 1 && w[1] == "list") {
                // property list <count type> <item type> <name>: a record's
                // size then depends on its data.
                p.list = true;
                p.type = "list";
                p.name = w.size() > 4 ? std::string(w[4]) : std::string();
                e.holds_list = true;
            } else {
                p.type = w.size() > 1 ? std::string(w[1]) : std::string();
                p.name = w.size() > 2 ? std::string(w[2]) : std::string();
                p.scalar = scalar_type(p.type);
                p.offset = (uint32_t)e.record_bytes;
                e.record_bytes += p.scalar.bytes;
            }
            e.props.push_back(std::move(p));
        }
        // comment, obj_info and anything else: skipped
    }
    return first ? Refusal{ATHX_INVALID_ARGUMENT, "not a PLY file"}
                 : Refusal{ATHX_INVALID_ARGUMENT, "PLY header has no end_header"};
}

// The vertex element's layout, checked as io::readSplatPly checks it.
Refusal vertex_layout(const Header& h, athx_ply_layout& L) {
    if (h.format != "binary_little_endian") {
        return {ATHX_UNSUPPORTED, "PLY format '" + h.format + "'; splat clouds are binary_little_endian"};
    }
    const Element* v = nullptr;
    for (const Element& e : h.elements) {
        if (e.name == "vertex") {
            v = &e;
            break;
        }
    }
    if (v == nullptr || v->count == 0) {
        return {ATHX_INVALID_ARGUMENT, "no vertices"};
    }
    if (v->holds_list) {
        return {ATHX_UNSUPPORTED, "a list in the vertex element"};
    }
    if (v->count > 0xFFFFFFFFull) {
        return {ATHX_UNSUPPORTED, "more than 2^32 splats"};
    }
    // The vertex records start after every earlier element's records, which
    // can be skipped only when their size does not depend on the data.
    uint64_t start = h.data_start;
    for (const Element* e = h.elements.data(); e != v; ++e) {
        if (e->holds_list) {
            return {ATHX_UNSUPPORTED, "a list element ('" + e->name + "') before 'vertex' makes its offset unknowable"};
        }
        start += e->record_bytes * e->count;
    }
    static const char* const kNeeded[] = {"x",       "y",     "z",     "opacity", "scale_0", "scale_1", "scale_2",
                                          "rot_0",   "rot_1", "rot_2", "rot_3",   "f_dc_0",  "f_dc_1",  "f_dc_2"};
    for (const char* name : kNeeded) {
        if (v->index_of(name) < 0) {
            return {ATHX_INVALID_ARGUMENT, std::string("no '") + name + "' -- this does not look like a trained splat cloud"};
        }
    }
    if (v->props.size() > ATHX_PLY_MAX_PROPERTIES) {
        return {ATHX_UNSUPPORTED, "more than " + std::to_string(ATHX_PLY_MAX_PROPERTIES) + " vertex properties"};
    }
    std::memset(&L, 0, sizeof L);
    L.data_offset = start;
    L.count = v->count;
    L.stride = (uint32_t)v->record_bytes;
    L.properties = (uint32_t)v->props.size();
    L.x = (uint32_t)v->index_of("x");
    L.y = (uint32_t)v->index_of("y");
    L.z = (uint32_t)v->index_of("z");
    L.opacity = (uint32_t)v->index_of("opacity");
    const char* const scale[3] = {"scale_0", "scale_1", "scale_2"};
    const char* const dc[3] = {"f_dc_0", "f_dc_1", "f_dc_2"};
    const char* const rot[4] = {"rot_0", "rot_1", "rot_2", "rot_3"};
    for (int k = 0; k < 3; ++k) {
        L.scale[k] = (uint32_t)v->index_of(scale[k]);
        L.dc[k] = (uint32_t)v->index_of(dc[k]);
    }
    for (int k = 0; k < 4; ++k) {
        L.rot[k] = (uint32_t)v->index_of(rot[k]);
    }
    // Rest harmonics f_rest_<j>: a colour's coefficients together (red's,
    // then green's, then blue's), all of them in one run of properties. The
    // file's count a colour must be a whole degree's 3, 8 or 15; fewer than
    // 9 in all are no harmonics.
    uint32_t rest = 0;
    for (const Property& p : v->props) {
        rest += p.name.compare(0, 7, "f_rest_") == 0 ? 1u : 0u;
    }
    const uint32_t per = rest >= 45 ? 15u : rest >= 24 ? 8u : rest >= 9 ? 3u : 0u;
    if (per != 0) {
        const int first = v->index_of("f_rest_0");
        const uint32_t in_file = rest / 3;
        for (uint32_t c = 0; c < 3; ++c) {
            for (uint32_t j = 0; j < per; ++j) {
                const uint32_t want = c * in_file + j;
                if (first < 0 || v->index_of("f_rest_" + std::to_string(want)) != first + (int)want) {
                    return {ATHX_UNSUPPORTED, "harmonics are not contiguous in the vertex element"};
                }
            }
        }
        if (in_file != per) {
            return {ATHX_UNSUPPORTED, std::to_string(rest) + " harmonic coefficients is not a whole degree"};
        }
        L.rest_base = (uint32_t)first;
        L.rest_per_colour = per;
    }
    for (size_t k = 0; k < v->props.size(); ++k) {
        const Property& p = v->props[k];
        if (p.scalar.code < 0) {
            return {ATHX_UNSUPPORTED, "property '" + p.name + "' of type '" + p.type + "'"};
        }
        L.offset[k] = p.offset;
        L.type[k] = (uint32_t)p.scalar.code;
    }
    return {};
}

athx_status report(const Refusal& r, char* message, uint64_t message_bytes) {
    if (message && message_bytes > 0) {
        const size_t n = r.reason.size() < message_bytes - 1 ? r.reason.size() : (size_t)(message_bytes - 1);
        std::memcpy(message, r.reason.data(), n);
        message[n] = '\0';
    }
    return r.status;
}

}  // namespace

extern "C" ATHX_API athx_status athx_ply_parse(const char* header, uint64_t bytes, athx_ply_layout* out,
                                               char* message, uint64_t message_bytes) try {
    if (message && message_bytes > 0) {
        message[0] = '\0';
    }
    if (!header || !out) {
        return report({ATHX_INVALID_ARGUMENT, "null header or layout"}, message, message_bytes);
    }
    Header h;
    Refusal r = read_header(std::string_view(header, bytes), h);
    if (r.status != ATHX_OK) {
        return report(r, message, message_bytes);
    }
    athx_ply_layout L{};
    r = vertex_layout(h, L);
    if (r.status != ATHX_OK) {
        return report(r, message, message_bytes);
    }
    *out = L;
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return report({ATHX_OUT_OF_MEMORY, "host allocation failed"}, message, message_bytes);
} catch (const std::exception& e) {
    return report({ATHX_INTERNAL_ERROR, e.what()}, message, message_bytes);
}

extern "C" ATHX_API athx_status athx_splat_layout(uint64_t bytes, athx_ply_layout* out, char* message,
                                                  uint64_t message_bytes) try {
    if (message && message_bytes > 0) {
        message[0] = '\0';
    }
    if (!out) {
        return report({ATHX_INVALID_ARGUMENT, "null layout"}, message, message_bytes);
    }
    // SplatReaders.cpp:208-213: whole 32-byte records, at least one.
    constexpr uint64_t kRecord = 32;
    if (bytes == 0 || bytes % kRecord != 0) {
        return report({ATHX_INVALID_ARGUMENT,
                       std::to_string(bytes) + " bytes is not a whole number of 32-byte .splat records"},
                      message, message_bytes);
    }
    if (bytes / kRecord > 0xFFFFFFFFull) {
        return report({ATHX_UNSUPPORTED, "more than 2^32 records"}, message, message_bytes);
    }
    athx_ply_layout L{};
    L.data_offset = 0;
    L.count = bytes / kRecord;
    L.stride = (uint32_t)kRecord;
    // pos f32 x 3 | scale f32 x 3 | rgba u8 x 4 | rot u8 x 4 (w x y z): six
    // floats, then eight bytes (SplatReaders.cpp:215-235).
    L.properties = 14;
    for (uint32_t k = 0; k < 14; ++k) {
        L.offset[k] = k < 6 ? 4 * k : 24 + (k - 6);
        L.type[k] = k < 6 ? ATHX_PLY_F32 : ATHX_PLY_U8;
    }
    L.x = 0;
    L.y = 1;
    L.z = 2;
    L.scale[0] = 3;
    L.scale[1] = 4;
    L.scale[2] = 5;
    L.dc[0] = 6;
    L.dc[1] = 7;
    L.dc[2] = 8;
    L.opacity = 9;
    L.rot[0] = 10;
    L.rot[1] = 11;
    L.rot[2] = 12;
    L.rot[3] = 13;
    L.rest_base = 0;
    L.rest_per_colour = 0;
    L.opacity_mode = 2;   // a byte
    L.scale_mode = 1;     // linear
    L.colour_mode = 2;    // a byte
    L.rotation_mode = 1;  // bytes
    *out = L;
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return report({ATHX_OUT_OF_MEMORY, "host allocation failed"}, message, message_bytes);
} catch (const std::exception& e) {
    return report({ATHX_INTERNAL_ERROR, e.what()}, message, message_bytes);
}
