// Scene XML loader: a small XML reader plus the tag semantics of the
// reference's loadFromXML (src/core/parser/parser.cpp:19-387): object tags
// are instantiated by `type` through the ObjectFactory with the properties of
// their child tags, children are attached with addChild/setParent, then the
// object is activated; `id` registers an object for later <ref id=".."/>;
// <transform> accumulates translate / scale / rotate / matrix / lookat
// operations, each applied on the left.  Errors are drawlab::Exception with
// the file name and line/column, as the reference reports them.
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

enum ETag {
    TScene = Object::EScene,
    TMesh = Object::EMesh,
    TBSDF = Object::EBSDF,
    TPhase = Object::EPhaseFunction,
    TEmitter = Object::EEmitter,
    TMedium = Object::EMedium,
    TCamera = Object::ECamera,
    TIntegrator = Object::EIntegrator,
    TSampler = Object::ESampler,
    TTest = Object::ETest,
    TRFilter = Object::EReconstructionFilter,
    TTexture = Object::ETexture,
    TBoolean = Object::EClassTypeCount,
    TInteger,
    TFloat,
    TString,
    TPoint,
    TVector,
    TColor,
    TTransform,
    TTranslate,
    TMatrix,
    TRotate,
    TScale,
    TLookAt,
    TRef,
    TInvalid
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

    static const std::map<std::string, ETag> tags = {
        {"scene", TScene},     {"mesh", TMesh},           {"bsdf", TBSDF},         {"emitter", TEmitter},
        {"camera", TCamera},   {"medium", TMedium},       {"phase", TPhase},       {"integrator", TIntegrator},
        {"sampler", TSampler}, {"rfilter", TRFilter},     {"test", TTest},         {"texture", TTexture},
        {"boolean", TBoolean}, {"integer", TInteger},     {"float", TFloat},       {"string", TString},
        {"point", TPoint},     {"vector", TVector},       {"color", TColor},       {"transform", TTransform},
        {"translate", TTranslate}, {"matrix", TMatrix},   {"rotate", TRotate},     {"scale", TScale},
        {"lookat", TLookAt},   {"ref", TRef}};

    auto check_attributes = [&](const XmlNode& node, std::set<std::string> attrs,
                                const std::set<std::string>& opt = {}) {
        for (const auto& a : node.attrs) {
            auto it = attrs.find(a.first);
            if (it == attrs.end()) {
                if (!opt.count(a.first))
                    throw Exception("Error while parsing \"%s\": unexpected attribute \"%s\" in \"%s\" at %s", filename,
                                    a.first, node.name, where(node.offset));
            } else {
                attrs.erase(it);
            }
        }
        if (!attrs.empty())
            throw Exception("Error while parsing \"%s\": missing attribute \"%s\" in \"%s\" at %s", filename,
                            *attrs.begin(), node.name, where(node.offset));
    };

    Transform transform;
    std::map<std::string, Object*> pool;

    std::function<Object*(const XmlNode&, PropertyList&, int)> parseTag = [&](const XmlNode& node, PropertyList& list,
                                                                            int parentTag) -> Object* {
        if (node.kind == XmlNode::Comment) return nullptr;
        if (node.kind != XmlNode::Element)
            throw Exception("Error while parsing \"%s\": unexpected content at %s", filename, where(node.offset));
        auto it = tags.find(node.name);
        if (it == tags.end())
            throw Exception("Error while parsing \"%s\": unexpected tag \"%s\" at %s", filename, node.name,
                            where(node.offset));
        const int tag = it->second;

        const bool hasParent = parentTag != TInvalid;
        const bool parentIsObject = hasParent && parentTag < Object::EClassTypeCount;
        const bool currentIsObject = tag < Object::EClassTypeCount || tag == TRef;
        const bool parentIsTransform = parentTag == TTransform;
        const bool currentIsTransformOp =
            tag == TTranslate || tag == TRotate || tag == TScale || tag == TLookAt || tag == TMatrix;

        if (!hasParent && !currentIsObject)
            throw Exception("Error while parsing \"%s\": root element \"%s\" must be a object (at %s)", filename,
                            node.name, where(node.offset));
        if (parentIsTransform != currentIsTransformOp)
            throw Exception("Error while parsing \"%s\": transform nodes can only contain transform operations (at %s)",
                            filename, where(node.offset));
        if (hasParent && !parentIsObject && !(parentIsTransform && currentIsTransformOp))
            throw Exception("Error while parsing \"%s\": node \"%s\" requires a object as parent (at %s)", filename,
                            node.name, where(node.offset));

        XmlNode self = node;
        if (tag == TScene && !self.has_attr("type")) self.attrs.emplace_back("type", "scene");
        if (tag == TTransform) transform.setIdentity();

        if (tag == TRef) {
            check_attributes(self, {"id"});
            const std::string id = self.attr("id");
            auto p = pool.find(id);
            if (p == pool.end())
                throw Exception("Error while parsing \"%s\": Can't find the referred node %s  (at %s)", filename, id,
                                where(node.offset));
            return p->second;
        }

        PropertyList props;
        std::vector<Object*> children;
        for (const XmlNode& ch : self.children) {
            Object* c = parseTag(ch, props, tag);
            if (c) children.push_back(c);
        }

        Object* result = nullptr;
        try {
            if (currentIsObject) {
                check_attributes(self, {"type"}, {"name", "id"});
                result = ObjectFactory::createInstance(self.attr("type"), props, self.attr("name"));
                if ((int)result->getClassType() != tag)
                    throw Exception("Unexpectedly constructed an object of type <%s> (expected type <%s>): %s",
                                    Object::classTypeName(result->getClassType()),
                                    Object::classTypeName((Object::EClassType)tag), result->toString());
                for (Object* c : children) {
                    result->addChild(c);
                    c->setParent(result);
                }
                result->activate();
                const std::string id = self.attr("id");
                if (!id.empty()) {
                    if (pool.count(id))
                        throw Exception("Error while parsing \"%s\": the node id has been used. (at %s)", filename,
                                        where(node.offset));
                    pool[id] = result;
                }
            } else {
                switch (tag) {
                    case TString:
                        check_attributes(self, {"name", "value"});
                        list.setString(self.attr("name"), self.attr("value"));
                        break;
                    case TFloat:
                        check_attributes(self, {"name", "value"});
                        list.setFloat(self.attr("name"), toFloat(self.attr("value")));
                        break;
                    case TInteger:
                        check_attributes(self, {"name", "value"});
                        list.setInteger(self.attr("name"), toInt(self.attr("value")));
                        break;
                    case TBoolean:
                        check_attributes(self, {"name", "value"});
                        list.setBoolean(self.attr("name"), toBool(self.attr("value")));
                        break;
                    case TPoint:
                        check_attributes(self, {"name", "value"});
                        list.setPoint(self.attr("name"), toVector3f(self.attr("value")));
                        break;
                    case TVector:
                        check_attributes(self, {"name", "value"});
                        list.setVector(self.attr("name"), toVector3f(self.attr("value")));
                        break;
                    case TColor:
                        check_attributes(self, {"name", "value"});
                        list.setColor(self.attr("name"), toVector3f(self.attr("value")));
                        break;
                    case TTransform:
                        check_attributes(self, {"name"});
                        list.setTransform(self.attr("name"), transform);
                        break;
                    case TTranslate:
                        check_attributes(self, {"value"});
                        transform = Transform(Matrix4f::translation(toVector3f(self.attr("value"))) * transform.matrix());
                        break;
                    case TMatrix: {
                        check_attributes(self, {"value"});
                        const std::vector<std::string> tok = tokenize(self.attr("value"), ", \t\r\n");
                        if (tok.size() != 16) throw Exception("Expected 16 values");
                        Matrix4f m;
                        for (int i = 0; i < 4; ++i)
                            for (int j = 0; j < 4; ++j) m.m[i][j] = toFloat(tok[i * 4 + j]);
                        transform = Transform(m * transform.matrix());
                    } break;
                    case TScale: {
                        check_attributes(self, {"value"});
                        // a single value is a uniform scale (the documented
                        // <scale value="5"/> form)
                        const std::vector<std::string> tok = tokenize(self.attr("value"));
                        const Vector3f v = tok.size() == 1 ? Vector3f(toFloat(tok[0])) : toVector3f(self.attr("value"));
                        transform = Transform(Matrix4f::scaling(v) * transform.matrix());
                    } break;
                    case TRotate: {
                        check_attributes(self, {"angle", "axis"});
                        const float angle = degToRad(toFloat(self.attr("angle")));
                        const Vector3f axis = toVector3f(self.attr("axis"));
                        transform = Transform(Matrix4f::rotation(axis, angle) * transform.matrix());
                    } break;
                    case TLookAt: {
                        check_attributes(self, {"origin", "target", "up"});
                        const Vector3f origin = toVector3f(self.attr("origin"));
                        const Vector3f target = toVector3f(self.attr("target"));
                        const Vector3f up = toVector3f(self.attr("up"));
                        const Vector3f dir = normalized(target - origin);
                        const Vector3f left = normalized(cross(normalized(up), dir));
                        const Vector3f newUp = normalized(cross(dir, left));
                        Matrix4f t;
                        for (int r = 0; r < 3; r++) {
                            t.m[r][0] = left[r];
                            t.m[r][1] = newUp[r];
                            t.m[r][2] = dir[r];
                            t.m[r][3] = origin[r];
                        }
                        t.m[3][3] = 1.f;
                        transform = Transform(t * transform.matrix());
                    } break;
                    default: throw Exception("Unhandled element \"%s\"", self.name);
                }
            }
        } catch (const Exception& e) {
            const std::string msg = e.what();
            if (msg.rfind("Error while parsing", 0) == 0) throw;
            throw Exception("Error while parsing \"%s\": %s (at %s)", filename, msg, where(node.offset));
        }
        return result;
    };

    PropertyList list;
    return parseTag(root, list, TInvalid);
}

}  // namespace drawlab
