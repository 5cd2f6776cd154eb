// Host-side object model of the DrawLab scene description, mirroring the
// reference's plugin interface (include/core/parser/object.h,
// include/core/parser/proplist.h, include/core/math/transform.h): scenes are
// trees of Objects built by name through the ObjectFactory from typed
// PropertyLists, exactly as the reference's XML loader builds them.  The HIP
// backend consumes the result through the C ABI (dl_hip.h); nothing here
// computes radiance.
#pragma once
#include <cstdint>
#include <cstdio>
#include <functional>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace drawlab {

// ---- errors (include/core/base/exception.h:8-16) ---------------------------
// The message is formatted printf-style: each %s / %i / %d / %f consumes the
// next argument, which is streamed with operator<<.
namespace detail {
inline void format_to(std::ostringstream& os, const char* f) {
    for (; *f; ++f) {
        if (f[0] == '%' && f[1] == '%') ++f;
        os << *f;
    }
}
template <typename T, typename... Rest>
void format_to(std::ostringstream& os, const char* f, const T& v, const Rest&... rest) {
    for (; *f; ++f) {
        if (f[0] == '%' && f[1] == '%') {
            os << '%';
            ++f;
            continue;
        }
        if (f[0] == '%' && f[1] != '\0') {
            os << v;
            format_to(os, f + 2, rest...);
            return;
        }
        os << *f;
    }
}
}  // namespace detail

template <typename... Args>
std::string format(const char* f, const Args&... args) {
    std::ostringstream os;
    detail::format_to(os, f, args...);
    return os.str();
}

// A warning on stderr (the reference logs these with spdlog::warn).
template <typename... Args>
void warn(const char* f, const Args&... args) {
    std::fprintf(stderr, "[warning] %s\n", format(f, args...).c_str());
}

class Exception : public std::runtime_error {
public:
    template <typename... Args>
    explicit Exception(const char* f, const Args&... args) : std::runtime_error(format(f, args...)) {}
};

// ---- small fixed-size math types ------------------------------------------
struct Vector3f {
    float v[3] = {0.f, 0.f, 0.f};
    Vector3f() = default;
    Vector3f(float x, float y, float z) : v{x, y, z} {}
    explicit Vector3f(float s) : v{s, s, s} {}
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    float z() const { return v[2]; }
};
using Point3f = Vector3f;
using Color3f = Vector3f;
using Normal3f = Vector3f;

inline std::ostream& operator<<(std::ostream& os, const Vector3f& a) {
    return os << "[" << a[0] << ", " << a[1] << ", " << a[2] << "]";
}

// Reference evaluation order: dot = ((0 + x x') + y y') + z z'
// (include/core/math/vector.h:329-336); length = sqrt(dot).
Vector3f operator+(const Vector3f& a, const Vector3f& b);
Vector3f operator-(const Vector3f& a, const Vector3f& b);
Vector3f operator*(const Vector3f& a, float s);
float dot(const Vector3f& a, const Vector3f& b);
Vector3f cross(const Vector3f& a, const Vector3f& b);
Vector3f normalized(const Vector3f& a);

// Row-major 4x4 matrix; products and matrix-vector products accumulate each
// entry as ((((0 + a0 b0) + a1 b1) + a2 b2) + a3 b3) like the reference's
// row(j).dot(col(i)) (include/core/math/matrix.h:216-224).
struct Matrix4f {
    float m[4][4] = {{0}};
    static Matrix4f identity();
    static Matrix4f translation(const Vector3f& t);
    static Matrix4f scaling(const Vector3f& s);
    static Matrix4f rotation(const Vector3f& axis, float angle_rad);
    Matrix4f operator*(const Matrix4f& b) const;
    void apply(const float in[4], float out[4]) const;
    Matrix4f transpose() const;
    Matrix4f inverse() const;  // cofactor expansion; throws on a singular matrix
};

// Homogeneous transform (include/core/math/transform.h:15-75).
class Transform {
public:
    Transform() : m_mat(Matrix4f::identity()) {}
    explicit Transform(const Matrix4f& m) : m_mat(m) {}
    const Matrix4f& matrix() const { return m_mat; }
    Point3f point(const Point3f& p) const;     // divides by w
    Vector3f vector(const Vector3f& v) const;  // w = 0
    Normal3f normal(const Normal3f& n) const;  // inverse transpose, w = 0
    void setIdentity() { m_mat = Matrix4f::identity(); }

private:
    Matrix4f m_mat;
};
inline std::ostream& operator<<(std::ostream& os, const Transform&) { return os << "Transform[...]"; }

// ---- string helpers (src/core/base/string.cpp:37-102) ----------------------
std::string toLower(const std::string& s);
bool toBool(const std::string& s);
int toInt(const std::string& s);
float toFloat(const std::string& s);
Vector3f toVector3f(const std::string& s);
std::vector<std::string> tokenize(const std::string& s, const std::string& delim = ", ", bool includeEmpty = false);
bool endsWith(const std::string& value, const std::string& ending);
float degToRad(float deg);

// ---- PropertyList (include/core/parser/proplist.h, src/core/parser/proplist.cpp)
class PropertyList {
public:
    enum Type { Boolean, Integer, Float, String, Color, Point, Vector, TransformT };

    void setBoolean(const std::string& name, bool v);
    void setInteger(const std::string& name, int v);
    void setFloat(const std::string& name, float v);
    void setString(const std::string& name, const std::string& v);
    void setColor(const std::string& name, const Color3f& v);
    void setPoint(const std::string& name, const Point3f& v);
    void setVector(const std::string& name, const Vector3f& v);
    void setTransform(const std::string& name, const Transform& v);

    // The one-argument getters throw "Property '<name>' is missing!" when the
    // property is absent; all getters throw "... has the wrong type!" when it
    // was declared with another tag.
    bool getBoolean(const std::string& name) const;
    bool getBoolean(const std::string& name, bool def) const;
    int getInteger(const std::string& name) const;
    int getInteger(const std::string& name, int def) const;
    float getFloat(const std::string& name) const;
    float getFloat(const std::string& name, float def) const;
    std::string getString(const std::string& name) const;
    std::string getString(const std::string& name, const std::string& def) const;
    Color3f getColor(const std::string& name) const;
    Color3f getColor(const std::string& name, const Color3f& def) const;
    Point3f getPoint(const std::string& name) const;
    Point3f getPoint(const std::string& name, const Point3f& def) const;
    Vector3f getVector(const std::string& name) const;
    Vector3f getVector(const std::string& name, const Vector3f& def) const;
    Transform getTransform(const std::string& name) const;
    Transform getTransform(const std::string& name, const Transform& def) const;

    bool has(const std::string& name) const { return m_props.count(name) != 0; }

private:
    struct Property {
        Type type = Boolean;
        bool b = false;
        int i = 0;
        float f = 0.f;
        std::string s;
        Vector3f vec;
        Transform t;
    };
    const Property* find(const std::string& name, Type type, bool required) const;
    Property& slot(const std::string& name, Type type);
    std::map<std::string, Property> m_props;
};

// ---- Object / ObjectFactory (include/core/parser/object.h:16-161) ----------
class Object {
public:
    enum EClassType {
        EScene = 0,
        EMesh,
        EBSDF,
        EPhaseFunction,
        EEmitter,
        EMedium,
        ECamera,
        EIntegrator,
        ESampler,
        ETest,
        EReconstructionFilter,
        ETexture,
        EClassTypeCount
    };
    virtual ~Object() {}
    virtual EClassType getClassType() const = 0;
    virtual void addChild(Object* child);
    virtual void setParent(Object*) {}
    virtual void activate() {}
    virtual std::string toString() const = 0;
    static std::string classTypeName(EClassType type);
    void setObjectName(const std::string& n) { m_object_name = n; }
    const std::string& getObjectName() const { return m_object_name; }

protected:
    std::string m_object_name;
};

class ObjectFactory {
public:
    using Constructor = std::function<Object*(const PropertyList&)>;
    static void registerClass(const std::string& name, const Constructor& c);
    // Throws 'A constructor for class "<name>" could not be found!'.
    static Object* createInstance(const std::string& name, const PropertyList& props,
                                  const std::string& object_name = "");
};

// Registration at static-initialisation time under the XML `type` name.
#define DRAWLAB_REGISTER_CLASS(cls, name)                                             \
    static Object* cls##_create(const PropertyList& list) { return new cls(list); } \
    static const bool cls##_registered = (ObjectFactory::registerClass(name, cls##_create), true);

// Loads a scene description (src/core/parser/parser.cpp:19-387).  Resources
// named by relative paths resolve against the XML file's directory.
Object* loadFromXML(const std::string& filename);

// Resolves a path against the directories registered while loading.
std::string resolvePath(const std::string& name);
void prependSearchPath(const std::string& dir);

}  // namespace drawlab
