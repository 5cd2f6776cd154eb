// Object model, property lists, math helpers and path resolution of the
// host plugin layer (mirrors src/core/parser/object.cpp, proplist.cpp,
// src/core/base/string.cpp, src/core/math/matrix.cpp).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iostream>

#include "drawlab/core.h"
#include "mat4.h"

namespace drawlab {

// ---- vector / matrix -------------------------------------------------------
Vector3f operator+(const Vector3f& a, const Vector3f& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
Vector3f operator-(const Vector3f& a, const Vector3f& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
Vector3f operator*(const Vector3f& a, float s) { return {a[0] * s, a[1] * s, a[2] * s}; }
float dot(const Vector3f& a, const Vector3f& b) {
    float s = 0.f;
    for (int i = 0; i < 3; i++) s += a[i] * b[i];
    return s;
}
Vector3f cross(const Vector3f& a, const Vector3f& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
Vector3f normalized(const Vector3f& a) {
    const float len = std::sqrt(dot(a, a));
    return {a[0] / len, a[1] / len, a[2] / len};
}

Matrix4f Matrix4f::identity() {
    Matrix4f r;
    for (int i = 0; i < 4; i++) r.m[i][i] = 1.f;
    return r;
}
Matrix4f Matrix4f::translation(const Vector3f& t) {
    Matrix4f r = identity();
    for (int i = 0; i < 3; i++) r.m[i][3] = t[i];
    return r;
}
Matrix4f Matrix4f::scaling(const Vector3f& s) {
    Matrix4f r;
    for (int i = 0; i < 3; i++) r.m[i][i] = s[i];
    r.m[3][3] = 1.f;
    return r;
}
// Rotation by `angle` about `axis` (Rodrigues' formula written out per entry,
// as Matrix4f::getRotateMatrix does; the axis is used as given).
Matrix4f Matrix4f::rotation(const Vector3f& axis, float angle) {
    const float x = axis[0], y = axis[1], z = axis[2];
    const float c = std::cos(angle), s = std::sin(angle), k = 1 - c;
    Matrix4f r;
    r.m[0][0] = c + k * x * x;
    r.m[0][1] = k * x * y - s * z;
    r.m[0][2] = k * x * z + s * y;
    r.m[1][0] = k * y * x + s * z;
    r.m[1][1] = c + k * y * y;
    r.m[1][2] = k * y * z - s * x;
    r.m[2][0] = k * z * x - s * y;
    r.m[2][1] = k * z * y + s * x;
    r.m[2][2] = c + k * z * z;
    r.m[3][3] = 1.f;
    return r;
}
Matrix4f Matrix4f::operator*(const Matrix4f& b) const {
    Matrix4f r;
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++) {
            float s = 0.f;
            for (int k = 0; k < 4; k++) s += m[j][k] * b.m[k][i];
            r.m[j][i] = s;
        }
    return r;
}
void Matrix4f::apply(const float in[4], float out[4]) const {
    for (int j = 0; j < 4; j++) {
        float s = 0.f;
        for (int k = 0; k < 4; k++) s += m[j][k] * in[k];
        out[j] = s;
    }
}
Matrix4f Matrix4f::transpose() const {
    Matrix4f r;
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++) r.m[j][i] = m[i][j];
    return r;
}
Matrix4f Matrix4f::inverse() const {
    Matrix4f r;
    if (!dlm::mat4_inverse(m, r.m)) throw Exception("Matrix is not invertible");
    return r;
}

Point3f Transform::point(const Point3f& p) const {
    const float in[4] = {p[0], p[1], p[2], 1.f};
    float out[4];
    m_mat.apply(in, out);
    return {out[0] / out[3], out[1] / out[3], out[2] / out[3]};
}
Vector3f Transform::vector(const Vector3f& v) const {
    const float in[4] = {v[0], v[1], v[2], 0.f};
    float out[4];
    m_mat.apply(in, out);
    return {out[0], out[1], out[2]};
}
Normal3f Transform::normal(const Normal3f& n) const {
    const float in[4] = {n[0], n[1], n[2], 0.f};
    float out[4];
    m_mat.inverse().transpose().apply(in, out);
    return {out[0], out[1], out[2]};
}

// ---- strings ---------------------------------------------------------------
std::string toLower(const std::string& s) {
    std::string r = s;
    std::transform(r.begin(), r.end(), r.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return r;
}
bool toBool(const std::string& s) {
    const std::string v = toLower(s);
    if (v == "false") return false;
    if (v == "true") return true;
    throw Exception("Could not parse boolean value \"%s\"", s);
}
int toInt(const std::string& s) {
    char* end = nullptr;
    const long r = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || *end != '\0') throw Exception("Could not parse integer value \"%s\"", s);
    return (int)r;
}
float toFloat(const std::string& s) {
    char* end = nullptr;
    const float r = std::strtof(s.c_str(), &end);
    if (end == s.c_str() || *end != '\0') throw Exception("Could not parse floating point value \"%s\"", s);
    return r;
}
std::vector<std::string> tokenize(const std::string& s, const std::string& delim, bool includeEmpty) {
    std::vector<std::string> out;
    size_t start = 0;
    while (true) {
        const size_t pos = s.find_first_of(delim, start);
        const std::string tok = s.substr(start, pos == std::string::npos ? std::string::npos : pos - start);
        if (!tok.empty() || includeEmpty) out.push_back(tok);
        if (pos == std::string::npos) break;
        start = pos + 1;
    }
    return out;
}
Vector3f toVector3f(const std::string& s) {
    const std::vector<std::string> t = tokenize(s);
    if (t.size() != 3) throw Exception("Expected 3 values");
    return {toFloat(t[0]), toFloat(t[1]), toFloat(t[2])};
}
bool endsWith(const std::string& value, const std::string& ending) {
    return ending.size() <= value.size() && std::equal(ending.rbegin(), ending.rend(), value.rbegin());
}
float degToRad(float deg) { return deg * (float)(M_PI / 180.0); }

// ---- PropertyList ----------------------------------------------------------
static const char* type_name(PropertyList::Type t) {
    static const char* n[] = {"boolean", "integer", "float", "string", "color", "point", "vector", "transform"};
    return n[t];
}
PropertyList::Property& PropertyList::slot(const std::string& name, Type type) {
    if (m_props.count(name)) std::cerr << "Property \"" << name << "\" was specified multiple times!" << std::endl;
    Property& p = m_props[name];
    p = Property();
    p.type = type;
    return p;
}
const PropertyList::Property* PropertyList::find(const std::string& name, Type type, bool required) const {
    auto it = m_props.find(name);
    if (it == m_props.end()) {
        if (required) throw Exception("Property '%s' is missing!", name);
        return nullptr;
    }
    if (it->second.type != type)
        throw Exception("Property '%s' has the wrong type! (expected <%s>)!", name, type_name(type));
    return &it->second;
}
void PropertyList::setBoolean(const std::string& n, bool v) { slot(n, Boolean).b = v; }
void PropertyList::setInteger(const std::string& n, int v) { slot(n, Integer).i = v; }
void PropertyList::setFloat(const std::string& n, float v) { slot(n, Float).f = v; }
void PropertyList::setString(const std::string& n, const std::string& v) { slot(n, String).s = v; }
void PropertyList::setColor(const std::string& n, const Color3f& v) { slot(n, Color).vec = v; }
void PropertyList::setPoint(const std::string& n, const Point3f& v) { slot(n, Point).vec = v; }
void PropertyList::setVector(const std::string& n, const Vector3f& v) { slot(n, Vector).vec = v; }
void PropertyList::setTransform(const std::string& n, const Transform& v) { slot(n, TransformT).t = v; }

bool PropertyList::getBoolean(const std::string& n) const { return find(n, Boolean, true)->b; }
bool PropertyList::getBoolean(const std::string& n, bool d) const {
    const Property* p = find(n, Boolean, false);
    return p ? p->b : d;
}
int PropertyList::getInteger(const std::string& n) const { return find(n, Integer, true)->i; }
int PropertyList::getInteger(const std::string& n, int d) const {
    const Property* p = find(n, Integer, false);
    return p ? p->i : d;
}
float PropertyList::getFloat(const std::string& n) const { return find(n, Float, true)->f; }
float PropertyList::getFloat(const std::string& n, float d) const {
    const Property* p = find(n, Float, false);
    return p ? p->f : d;
}
std::string PropertyList::getString(const std::string& n) const { return find(n, String, true)->s; }
std::string PropertyList::getString(const std::string& n, const std::string& d) const {
    const Property* p = find(n, String, false);
    return p ? p->s : d;
}
Color3f PropertyList::getColor(const std::string& n) const { return find(n, Color, true)->vec; }
Color3f PropertyList::getColor(const std::string& n, const Color3f& d) const {
    const Property* p = find(n, Color, false);
    return p ? p->vec : d;
}
Point3f PropertyList::getPoint(const std::string& n) const { return find(n, Point, true)->vec; }
Point3f PropertyList::getPoint(const std::string& n, const Point3f& d) const {
    const Property* p = find(n, Point, false);
    return p ? p->vec : d;
}
Vector3f PropertyList::getVector(const std::string& n) const { return find(n, Vector, true)->vec; }
Vector3f PropertyList::getVector(const std::string& n, const Vector3f& d) const {
    const Property* p = find(n, Vector, false);
    return p ? p->vec : d;
}
Transform PropertyList::getTransform(const std::string& n) const { return find(n, TransformT, true)->t; }
Transform PropertyList::getTransform(const std::string& n, const Transform& d) const {
    const Property* p = find(n, TransformT, false);
    return p ? p->t : d;
}

// ---- Object / factory ------------------------------------------------------
void Object::addChild(Object*) {
    throw Exception("Object::addChild() is not implemented for objects of type '%s'!",
                    classTypeName(getClassType()));
}
std::string Object::classTypeName(EClassType type) {
    switch (type) {
        case EScene: return "scene";
        case EMesh: return "mesh";
        case EBSDF: return "bsdf";
        case EEmitter: return "emitter";
        case ECamera: return "camera";
        case EIntegrator: return "integrator";
        case ESampler: return "sampler";
        case ETest: return "test";
        case ETexture: return "texture";
        case EReconstructionFilter: return "rfilter";
        default: return "<unknown>";
    }
}
static std::map<std::string, ObjectFactory::Constructor>& registry() {
    static std::map<std::string, ObjectFactory::Constructor> r;
    return r;
}
void ObjectFactory::registerClass(const std::string& name, const Constructor& c) { registry()[name] = c; }
Object* ObjectFactory::createInstance(const std::string& name, const PropertyList& props,
                                      const std::string& object_name) {
    auto it = registry().find(name);
    if (it == registry().end()) throw Exception("A constructor for class \"%s\" could not be found!", name);
    Object* o = it->second(props);
    o->setObjectName(object_name);
    return o;
}

// ---- file resolver ---------------------------------------------------------
static std::vector<std::string>& search_paths() {
    static std::vector<std::string> p;
    return p;
}
void prependSearchPath(const std::string& dir) { search_paths().insert(search_paths().begin(), dir); }
std::string resolvePath(const std::string& name) {
    if (!name.empty() && name[0] == '/') return name;
    for (const std::string& d : search_paths()) {
        const std::string p = d.empty() ? name : d + "/" + name;
        if (FILE* f = std::fopen(p.c_str(), "rb")) {
            std::fclose(f);
            return p;
        }
    }
    return name;
}

}  // namespace drawlab
