// Scene XML loader: a small XML reader and a table-driven loader for the
// scene format of the reference's loadFromXML (src/core/parser/parser.cpp:
// 19-387).  Each tag has a role (object, reference, property, <transform>,
// transform operation) with its attributes and handler in one table; where a
// role may appear is a second table.  Objects are instantiated by `type`
// through the ObjectFactory with the properties their children set, get
// their child objects, and are activated; `id` names an object for a later
// <ref id=".."/>; a <transform> is the product of its operations, each
// applied on the left.  The tag names and the error messages (with the file
// name and line/column) are the reference's, so scenes and their errors read
// the same.
#include <algorithm>
#include <cctype>
#include <cstring>
#include <cmath>
#include <fstream>
#include <functional>
#include <set>
#include <sstream>

#include "drawlab/core.h"

namespace drawlab {
namespace {

// ---- minimal XML DOM --------------------------------------------------------
struct XmlNode {
    enum Kind { Element, Comment, Text } kind = Element;
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<XmlNode> children;
    size_t offset = 0;

    const char* attr(const std::string& n) const {
        for (const auto& a : attrs)
            if (a.first == n) return a.second.c_str();
        return "";
    }
    bool has_attr(const std::string& n) const {
        for (const auto& a : attrs)
            if (a.first == n) return true;
        return false;
    }
};

class XmlReader {
public:
    XmlReader(const std::string& text, std::function<std::string(size_t)> where)
        : s_(text), where_(std::move(where)) {}

    XmlNode parse_document() {
        skip_misc();
        if (pos_ >= s_.size() || s_[pos_] != '<') fail("no root element");
        XmlNode root = parse_element();
        skip_misc();
        if (pos_ < s_.size()) fail("unexpected content after the root element");
        return root;
    }

private:
    [[noreturn]] void fail(const std::string& what) const { throw Exception("%s (at %s)", what, where_(pos_)); }

    bool starts(const char* p) const { return s_.compare(pos_, std::strlen(p), p) == 0; }
    void skip_ws() {
        while (pos_ < s_.size() && std::isspace((unsigned char)s_[pos_])) ++pos_;
    }
    void skip_until(const char* end) {
        const size_t e = s_.find(end, pos_);
        if (e == std::string::npos) fail(std::string("unterminated construct, expected ") + end);
        pos_ = e + std::strlen(end);
    }
    // whitespace, comments, processing instructions, DOCTYPE
    void skip_misc() {
        for (;;) {
            skip_ws();
            if (starts("<?"))
                skip_until("?>");
            else if (starts("<!--"))
                skip_until("-->");
            else if (starts("<!DOCTYPE"))
                skip_until(">");
            else
                return;
        }
    }
    std::string parse_name() {
        const size_t b = pos_;
        while (pos_ < s_.size() && (std::isalnum((unsigned char)s_[pos_]) || s_[pos_] == '_' || s_[pos_] == '-' ||
                                    s_[pos_] == ':' || s_[pos_] == '.'))
            ++pos_;
        if (b == pos_) fail("expected a name");
        return s_.substr(b, pos_ - b);
    }
    static std::string decode(const std::string& v) {
        std::string out;
        for (size_t i = 0; i < v.size(); ++i) {
            if (v[i] != '&') {
                out += v[i];
                continue;
            }
            const size_t e = v.find(';', i);
            const std::string ent = e == std::string::npos ? "" : v.substr(i + 1, e - i - 1);
            if (ent == "lt")
                out += '<';
            else if (ent == "gt")
                out += '>';
            else if (ent == "amp")
                out += '&';
            else if (ent == "quot")
                out += '"';
            else if (ent == "apos")
                out += '\'';
            else if (!ent.empty() && ent[0] == '#')
                out += (char)std::strtol(ent.c_str() + (ent[1] == 'x' ? 2 : 1), nullptr, ent[1] == 'x' ? 16 : 10);
            else {
                out += '&';
                continue;
            }
            i = e;
        }
        return out;
    }
    XmlNode parse_element() {
        XmlNode n;
        n.offset = pos_;
        ++pos_;  // '<'
        n.name = parse_name();
        for (;;) {
            skip_ws();
            if (pos_ >= s_.size()) fail("unterminated start tag");
            if (starts("/>")) {
                pos_ += 2;
                return n;
            }
            if (s_[pos_] == '>') {
                ++pos_;
                break;
            }
            const std::string an = parse_name();
            skip_ws();
            if (pos_ >= s_.size() || s_[pos_] != '=') fail("expected '=' after attribute name");
            ++pos_;
            skip_ws();
            if (pos_ >= s_.size() || (s_[pos_] != '"' && s_[pos_] != '\'')) fail("expected a quoted attribute value");
            const char q = s_[pos_++];
            const size_t e = s_.find(q, pos_);
            if (e == std::string::npos) fail("unterminated attribute value");
            n.attrs.emplace_back(an, decode(s_.substr(pos_, e - pos_)));
            pos_ = e + 1;
        }
        // content
        for (;;) {
            const size_t text_begin = pos_;
            while (pos_ < s_.size() && s_[pos_] != '<') ++pos_;
            if (pos_ >= s_.size()) fail("unterminated element <" + n.name + ">");
            for (size_t i = text_begin; i < pos_; ++i)
                if (!std::isspace((unsigned char)s_[i])) {
                    XmlNode t;
                    t.kind = XmlNode::Text;
                    t.offset = i;
                    n.children.push_back(t);
                    break;
                }
            if (starts("</")) {
                pos_ += 2;
                const std::string close = parse_name();
                if (close != n.name) fail("mismatched end tag </" + close + "> for <" + n.name + ">");
                skip_ws();
                if (pos_ >= s_.size() || s_[pos_] != '>') fail("expected '>'");
                ++pos_;
                return n;
            }
            if (starts("<!--")) {
                XmlNode c;
                c.kind = XmlNode::Comment;
                c.offset = pos_;
                skip_until("-->");
                n.children.push_back(c);
                continue;
            }
            if (starts("<?")) {
                skip_until("?>");
                continue;
            }
            n.children.push_back(parse_element());
        }
    }

    const std::string& s_;
    std::function<std::string(size_t)> where_;
    size_t pos_ = 0;
};

// ---- what each tag is ------------------------------------------------------
// Every element plays one role.  Objects are built by the factory from the
// properties their children set; a reference names an object built earlier;
// a property sets one entry of its parent object's PropertyList; a
// <transform> property is the product of the transform operations inside it.
enum class Role { Object, Reference, Property, Transform, TransformOp };

// Where an element sits: at the root, inside an object, inside a
// <transform>, or inside any other element (a property or an operation).
enum class Place { Root, InObject, InTransform, InLeaf };

using PropertySetter = void (*)(PropertyList&, const std::string& name, const std::string& value);
using OpMatrix = Matrix4f (*)(const XmlNode&);

struct TagSpec {
    Role role;
    Object::EClassType cls;                 // Role::Object
    std::vector<std::string> required;      // attributes
    std::vector<std::string> optional;
    PropertySetter set = nullptr;           // Role::Property
    OpMatrix op = nullptr;                  // Role::TransformOp
};

Vector3f uniform_or_vector(const std::string& v) {
    // <scale value="s"/> is a uniform scale, as the documented form reads
    const std::vector<std::string> tok = tokenize(v);
    return tok.size() == 1 ? Vector3f(toFloat(tok[0])) : toVector3f(v);
}

// Camera-to-world of a camera at `origin` looking at `target`: columns are
// the camera's x axis (up x forward), its y axis (forward x x) and forward.
Matrix4f look_at(const XmlNode& n) {
    const Vector3f eye = toVector3f(n.attr("origin"));
    const Vector3f fwd = normalized(toVector3f(n.attr("target")) - eye);
    const Vector3f xaxis = normalized(cross(normalized(toVector3f(n.attr("up"))), fwd));
    const Vector3f yaxis = normalized(cross(fwd, xaxis));
    Matrix4f m;
    const Vector3f* cols[4] = {&xaxis, &yaxis, &fwd, &eye};
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 3; r++) m.m[r][c] = (*cols[c])[r];
    m.m[3][3] = 1.f;
    return m;
}

Matrix4f matrix_of(const XmlNode& n) {
    const std::vector<std::string> tok = tokenize(n.attr("value"), ", \t\r\n");
    if (tok.size() != 16) throw Exception("Expected 16 values");
    Matrix4f m;
    for (size_t k = 0; k < 16; ++k) m.m[k / 4][k % 4] = toFloat(tok[k]);
    return m;
}

const std::map<std::string, TagSpec>& tag_table() {
    static const std::map<std::string, TagSpec> t = [] {
        std::map<std::string, TagSpec> m;
        const std::vector<std::string> obj_req = {"type"}, obj_opt = {"name", "id"}, nv = {"name", "value"};
        const std::pair<const char*, Object::EClassType> objects[] = {
            {"scene", Object::EScene},          {"mesh", Object::EMesh},        {"bsdf", Object::EBSDF},
            {"emitter", Object::EEmitter},      {"camera", Object::ECamera},    {"medium", Object::EMedium},
            {"phase", Object::EPhaseFunction},  {"integrator", Object::EIntegrator},
            {"sampler", Object::ESampler},      {"rfilter", Object::EReconstructionFilter},
            {"test", Object::ETest},            {"texture", Object::ETexture}};
        for (const auto& o : objects) m[o.first] = TagSpec{Role::Object, o.second, obj_req, obj_opt};
        m["ref"] = TagSpec{Role::Reference, Object::EClassTypeCount, {"id"}, {}};
        auto prop = [&](const char* tag, PropertySetter f) {
            m[tag] = TagSpec{Role::Property, Object::EClassTypeCount, nv, {}, f};
        };
        prop("string", [](PropertyList& l, const std::string& k, const std::string& v) { l.setString(k, v); });
        prop("float", [](PropertyList& l, const std::string& k, const std::string& v) { l.setFloat(k, toFloat(v)); });
        prop("integer", [](PropertyList& l, const std::string& k, const std::string& v) { l.setInteger(k, toInt(v)); });
        prop("boolean", [](PropertyList& l, const std::string& k, const std::string& v) { l.setBoolean(k, toBool(v)); });
        prop("point", [](PropertyList& l, const std::string& k, const std::string& v) { l.setPoint(k, toVector3f(v)); });
        prop("vector", [](PropertyList& l, const std::string& k, const std::string& v) { l.setVector(k, toVector3f(v)); });
        prop("color", [](PropertyList& l, const std::string& k, const std::string& v) { l.setColor(k, toVector3f(v)); });
        m["transform"] = TagSpec{Role::Transform, Object::EClassTypeCount, {"name"}, {}};
        auto op = [&](const char* tag, std::vector<std::string> req, OpMatrix f) {
            m[tag] = TagSpec{Role::TransformOp, Object::EClassTypeCount, std::move(req), {}, nullptr, f};
        };
        op("translate", {"value"}, [](const XmlNode& n) { return Matrix4f::translation(toVector3f(n.attr("value"))); });
        op("scale", {"value"}, [](const XmlNode& n) { return Matrix4f::scaling(uniform_or_vector(n.attr("value"))); });
        op("rotate", {"angle", "axis"}, [](const XmlNode& n) {
            return Matrix4f::rotation(toVector3f(n.attr("axis")), degToRad(toFloat(n.attr("angle"))));
        });
        op("matrix", {"value"}, matrix_of);
        op("lookat", {"origin", "target", "up"}, look_at);
        return m;
    }();
    return t;
}

// Which roles may appear where, and the message for those that may not.
enum class Misplaced { No, RootNotObject, OnlyTransformOps, NeedsObjectParent };
Misplaced placement(Place where, Role role) {
    static const Misplaced ok = Misplaced::No, root = Misplaced::RootNotObject, ops = Misplaced::OnlyTransformOps,
                           obj = Misplaced::NeedsObjectParent;
    //                          Object Reference Property Transform TransformOp
    static const Misplaced table[4][5] = {{ok, ok, root, root, root},   // Root
                                          {ok, ok, ok, ok, ops},        // InObject
                                          {ops, ops, ops, ops, ok},     // InTransform
                                          {obj, obj, obj, obj, ops}};   // InLeaf
    return table[(int)where][(int)role];
}

class SceneLoader {
public:
    SceneLoader(const std::string& filename, std::function<std::string(size_t)> where)
        : file_(filename), where_(std::move(where)) {}

    Object* load(const XmlNode& root) {
        PropertyList unused;
        return visit(root, Place::Root, unused, nullptr);
    }

private:
    template <typename... Args>
    Exception error(const char* fmt, const Args&... args) const {
        return Exception(("Error while parsing \"%s\": " + std::string(fmt)).c_str(), file_, args...);
    }

    void check_attributes(const XmlNode& n, const TagSpec& spec) const {
        std::set<std::string> missing(spec.required.begin(), spec.required.end());
        for (const auto& a : n.attrs) {
            if (missing.erase(a.first)) continue;
            if (std::find(spec.optional.begin(), spec.optional.end(), a.first) != spec.optional.end()) continue;
            throw error("unexpected attribute \"%s\" in \"%s\" at %s", a.first, n.name, where_(n.offset));
        }
        if (!missing.empty())
            throw error("missing attribute \"%s\" in \"%s\" at %s", *missing.begin(), n.name, where_(n.offset));
    }

    // Parses `node` found at `where`; a property lands in `props`, an
    // operation multiplies `*tf`; objects and references are returned.
    Object* visit(const XmlNode& node, Place where, PropertyList& props, Matrix4f* tf) {
        if (node.kind == XmlNode::Comment) return nullptr;
        if (node.kind != XmlNode::Element) throw error("unexpected content at %s", where_(node.offset));
        const auto& table = tag_table();
        const auto it = table.find(node.name);
        if (it == table.end()) throw error("unexpected tag \"%s\" at %s", node.name, where_(node.offset));
        const TagSpec& spec = it->second;
        switch (placement(where, spec.role)) {
            case Misplaced::RootNotObject:
                throw error("root element \"%s\" must be a object (at %s)", node.name, where_(node.offset));
            case Misplaced::OnlyTransformOps:
                throw error("transform nodes can only contain transform operations (at %s)", where_(node.offset));
            case Misplaced::NeedsObjectParent:
                throw error("node \"%s\" requires a object as parent (at %s)", node.name, where_(node.offset));
            case Misplaced::No: break;
        }
        if (spec.role == Role::Reference) {
            check_attributes(node, spec);
            const auto found = ids_.find(node.attr("id"));
            if (found == ids_.end())
                throw error("Can't find the referred node %s  (at %s)", node.attr("id"), where_(node.offset));
            return found->second;
        }
        try {
            switch (spec.role) {
                case Role::Object: return build_object(node, spec);
                case Role::Property:
                    visit_children(node, Place::InLeaf, props, nullptr);
                    check_attributes(node, spec);
                    spec.set(props, node.attr("name"), node.attr("value"));
                    return nullptr;
                case Role::Transform: {
                    Matrix4f product = Matrix4f::identity();  // each operation multiplies on the left
                    visit_children(node, Place::InTransform, props, &product);
                    check_attributes(node, spec);
                    props.setTransform(node.attr("name"), Transform(product));
                    return nullptr;
                }
                case Role::TransformOp:
                    visit_children(node, Place::InLeaf, props, nullptr);
                    check_attributes(node, spec);
                    *tf = spec.op(node) * *tf;
                    return nullptr;
                case Role::Reference: break;
            }
        } catch (const Exception& e) {
            const std::string msg = e.what();
            if (msg.rfind("Error while parsing", 0) == 0) throw;
            throw error("%s (at %s)", msg, where_(node.offset));
        }
        return nullptr;
    }

    std::vector<Object*> visit_children(const XmlNode& node, Place where, PropertyList& props, Matrix4f* tf) {
        std::vector<Object*> objs;
        for (const XmlNode& ch : node.children)
            if (Object* o = visit(ch, where, props, tf)) objs.push_back(o);
        return objs;
    }

    Object* build_object(const XmlNode& node, const TagSpec& spec) {
        PropertyList props;
        const std::vector<Object*> kids = visit_children(node, Place::InObject, props, nullptr);
        XmlNode decl = node;
        if (spec.cls == Object::EScene && !decl.has_attr("type")) decl.attrs.emplace_back("type", "scene");
        check_attributes(decl, spec);
        Object* obj = ObjectFactory::createInstance(decl.attr("type"), props, decl.attr("name"));
        if (obj->getClassType() != spec.cls)
            throw Exception("Unexpectedly constructed an object of type <%s> (expected type <%s>): %s",
                            Object::classTypeName(obj->getClassType()), Object::classTypeName(spec.cls),
                            obj->toString());
        for (Object* k : kids) {
            obj->addChild(k);
            k->setParent(obj);
        }
        obj->activate();
        const std::string id = decl.attr("id");
        if (!id.empty() && !ids_.emplace(id, obj).second)
            throw error("the node id has been used. (at %s)", where_(node.offset));
        return obj;
    }

    std::string file_;
    std::function<std::string(size_t)> where_;
    std::map<std::string, Object*> ids_;  // objects with an id, for <ref>
};

}  // namespace

Object* loadFromXML(const std::string& filename) {
    std::ifstream in(filename, std::ios::binary);
    if (!in) throw Exception("Error while parsing \"%s\": file could not be opened", filename);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();

    const size_t slash = filename.find_last_of('/');
    prependSearchPath(slash == std::string::npos ? std::string(".") : filename.substr(0, slash));

    auto where = [&](size_t pos) -> std::string {
        int line = 1;
        size_t line_start = 0;
        for (size_t i = 0; i < pos && i < text.size(); ++i)
            if (text[i] == '\n') {
                ++line;
                line_start = i + 1;
            }
        return format("line %i, col %i", line, (int)(pos - line_start) + 1);
    };

    XmlNode root;
    try {
        root = XmlReader(text, where).parse_document();
    } catch (const Exception& e) {
        throw Exception("Error while parsing \"%s\": %s", filename, e.what());
    }
    return SceneLoader(filename, where).load(root);
}

}  // namespace drawlab
