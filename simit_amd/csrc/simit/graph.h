// simit::Set / FieldRef / ElementRef work-alike: the host-side data model
// that Simit programs bind (reference: src/graph.h:39-1072, src/graph.cpp).
//
// Sets hold structure-of-arrays fields; an edge set also holds its endpoints
// as int32[size * cardinality] in element order (graph.h:105, :529, :663-679).
// Elements are added one at a time and are numbered 0, 1, 2, ... in the
// order they were added.  The HIP backend borrows these arrays at bind time
// and copies them to the device at init / unmapArgs.
#pragma once

#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace simit {

// Errors throw, as the reference's simit_uassert/simit_ierror do
// (src/error.cpp:87-97).
class SimitException : public std::runtime_error {
 public:
  explicit SimitException(const std::string &m) : std::runtime_error(m) {}
};

class ElementRef {
 public:
  ElementRef() : ident(-1) {}
  explicit ElementRef(int i) : ident(i) {}
  int getIdent() const { return ident; }
  bool defined() const { return ident >= 0; }
  bool operator==(const ElementRef &o) const { return ident == o.ident; }
  bool operator!=(const ElementRef &o) const { return ident != o.ident; }
  bool operator<(const ElementRef &o) const { return ident < o.ident; }

 private:
  int ident;
};

enum class FieldType { Float64, Float32, Int32, Bool };

template <typename T> struct FieldTypeOf;
template <> struct FieldTypeOf<double> { static constexpr FieldType value = FieldType::Float64; };
template <> struct FieldTypeOf<float> { static constexpr FieldType value = FieldType::Float32; };
template <> struct FieldTypeOf<int> { static constexpr FieldType value = FieldType::Int32; };
template <> struct FieldTypeOf<bool> { static constexpr FieldType value = FieldType::Bool; };

inline size_t fieldTypeSize(FieldType t) {
  switch (t) {
    case FieldType::Float64: return 8;
    case FieldType::Float32: return 4;
    case FieldType::Int32: return 4;
    case FieldType::Bool: return 1;
  }
  return 0;
}

// One field of a set: size() elements of ncomp components of type.
struct FieldData {
  std::string name;
  FieldType type;
  int ncomp;
  std::vector<unsigned char> bytes;
  size_t elementBytes() const { return fieldTypeSize(type) * (size_t)ncomp; }
  void *data() { return bytes.data(); }
  const void *data() const { return bytes.data(); }
};

class Set;

// A vector (or matrix) element of a field, by reference.
template <typename T>
class TensorRef {
 public:
  TensorRef(T *p, int n) : p(p), n(n) {}
  T &operator()(int i) const { return p[i]; }
  T &operator()(int i, int j, int cols) const { return p[i * cols + j]; }
  TensorRef &operator=(std::initializer_list<T> v) {
    int i = 0;
    for (const T &x : v) {
      if (i >= n) throw SimitException("too many values for a field element");
      p[i++] = x;
    }
    return *this;
  }
  int size() const { return n; }
  T *data() const { return p; }

 private:
  T *p;
  int n;
};

// Typed handle to a field (graph.h:755-927).  dims = () for scalars.
template <typename T, int... dims>
class FieldRef {
 public:
  static constexpr int kComp = sizeof...(dims) == 0 ? 1 : (1 * ... * dims);
  FieldRef() : owner(nullptr), field(nullptr) {}
  FieldRef(Set *s, FieldData *f) : owner(s), field(f) {}

  T *ptr(ElementRef e) const {
    return reinterpret_cast<T *>(field->data()) + (size_t)e.getIdent() * kComp;
  }
  // scalars: fld(e) is a T&; vectors: fld(e) is a TensorRef
  template <int K = sizeof...(dims)>
  typename std::enable_if<K == 0, T &>::type operator()(ElementRef e) const { return *ptr(e); }
  template <int K = sizeof...(dims)>
  typename std::enable_if<(K > 0), TensorRef<T>>::type operator()(ElementRef e) const {
    return TensorRef<T>(ptr(e), kComp);
  }
  template <int K = sizeof...(dims)>
  typename std::enable_if<K == 0, T>::type get(ElementRef e) const { return *ptr(e); }
  template <int K = sizeof...(dims)>
  typename std::enable_if<(K > 0), TensorRef<T>>::type get(ElementRef e) const {
    return TensorRef<T>(ptr(e), kComp);
  }
  void set(ElementRef e, T v) const { *ptr(e) = v; }
  void set(ElementRef e, std::initializer_list<T> v) const { TensorRef<T>(ptr(e), kComp) = v; }
  bool defined() const { return field != nullptr; }

 private:
  Set *owner;
  FieldData *field;
};

class Set {
 public:
  Set() : Set(std::string()) {}
  explicit Set(const std::string &name) : name(name) {}
  // An edge set whose endpoints are elements of the given sets (graph.h:101-118).
  template <typename... Sets>
  Set(const Set &first, const Sets &...rest) : Set(std::string(), first, rest...) {}
  template <typename... Sets>
  Set(const std::string &name, const Set &first, const Sets &...rest) : name(name) {
    endpointSets = {&first, &rest...};
  }
  // A grid edge set over `points` (graph.h:121-168): points must be empty
  // and receives d0*d1*... points, numbered with the first coordinate
  // fastest; this set gets one edge per point and axis, edge p*ndims + k
  // joining point p to its periodic neighbour along axis k (the numbering
  // getGridEdge returns).
  Set(Set &points, std::vector<int> dims) : Set(std::string(), points, std::move(dims)) {}
  Set(const std::string &name, Set &points, std::vector<int> dims) : name(name) {
    if (dims.empty())
      throw SimitException("Grid Edge Set constructor takes the underlying point set and a "
                           "vector of integer dimension sizes");
    if (points.getSize() != 0)
      throw SimitException("Grid Edge Set constructor must be passed an empty underlying point set");
    for (int d : dims)
      if (d <= 0) throw SimitException("grid dimensions must be positive");
    kind = Grid;
    dimensions = dims;
    endpointSets = {&points, &points};
    size_t total = 1;
    for (int d : dims) total *= (size_t)d;
    for (size_t p = 0; p < total; ++p) points.add();
    const int nd = (int)dims.size();
    for (size_t p = 0; p < total; ++p)
      for (int k = 0; k < nd; ++k) {
        // the neighbour of p one step along axis k, wrapping around
        size_t stride = 1;
        for (int i = 0; i < k; ++i) stride *= (size_t)dims[i];
        const size_t c = (p / stride) % (size_t)dims[k];
        const size_t q = c + 1 == (size_t)dims[k] ? p - c * stride : p + stride;
        add(ElementRef((int)p), ElementRef((int)q));
      }
  }
  Set(const Set &) = delete;
  Set &operator=(const Set &) = delete;

  enum Kind { Unstructured, Grid };
  Kind getKind() const { return kind; }
  // a grid edge set's dimension sizes
  const std::vector<int> &getDimensions() const {
    if (kind != Grid) throw SimitException("Can only retrieve dimensions for a grid edge set");
    return dimensions;
  }
  // the grid point at coords (coords[0] fastest), graph.h:191-209
  ElementRef getGridPoint(const std::vector<int> &coords) const {
    return ElementRef((int)gridIndex(coords, "grid point"));
  }
  // the grid edge leaving coords along axis dir, graph.h:212-233
  ElementRef getGridEdge(const std::vector<int> &coords, int dir) const {
    const int64_t nd = (int64_t)dimensions.size();
    const int64_t index = gridIndex(coords, "grid edge") * nd + dir;
    if (dir < 0 || dir >= nd)
      throw SimitException("Coordinates must not be negative and must fall within the grid "
                           "dimensions");
    return ElementRef((int)index);
  }
  std::vector<int> getGridPointCoords(ElementRef elt) const {
    if (kind != Grid) throw SimitException("Cannot retrieve grid point coords of non-grid set");
    std::vector<int> c;
    int64_t i = elt.getIdent();
    for (int d : dimensions) {
      c.push_back((int)(i % d));
      i /= d;
    }
    return c;
  }

  const std::string &getName() const { return name; }
  int getSize() const { return numElements; }
  int getCardinality() const { return (int)endpointSets.size(); }
  const Set *getEndpointSet(int i) const { return endpointSets.at(i); }

  // vertex add (graph.h:295-307)
  ElementRef add() {
    if (getCardinality() != 0)
      throw SimitException("Set::add() without endpoints on an edge set");
    return grow();
  }
  // edge add: one ElementRef per endpoint set
  template <typename... E>
  ElementRef add(ElementRef first, E... rest) {
    std::vector<ElementRef> eps{first, rest...};
    if ((int)eps.size() != getCardinality())
      throw SimitException("Set::add: wrong number of endpoints for " + name);
    for (int i = 0; i < getCardinality(); ++i)
      if (eps[i].getIdent() < 0 || eps[i].getIdent() >= endpointSets[i]->getSize())
        throw SimitException("Set::add: endpoint " + std::to_string(eps[i].getIdent()) +
                             " out of range");
    ElementRef r = grow();
    for (int i = 0; i < getCardinality(); ++i) endpoints.push_back(eps[i].getIdent());
    return r;
  }

  template <typename T, int... dims>
  FieldRef<T, dims...> addField(const std::string &fname) {
    if (fields.count(fname)) throw SimitException("field " + fname + " already exists");
    auto f = std::make_unique<FieldData>();
    f->name = fname;
    f->type = FieldTypeOf<T>::value;
    f->ncomp = FieldRef<T, dims...>::kComp;
    f->bytes.assign(f->elementBytes() * (size_t)numElements, 0);
    FieldData *raw = f.get();
    fields[fname] = std::move(f);
    order.push_back(fname);
    return FieldRef<T, dims...>(this, raw);
  }
  template <typename T, int... dims>
  FieldRef<T, dims...> getField(const std::string &fname) {
    FieldData *f = field(fname);
    if (!f || f->type != FieldTypeOf<T>::value || f->ncomp != FieldRef<T, dims...>::kComp)
      throw SimitException("no field " + fname + " of that type in set " + name);
    return FieldRef<T, dims...>(this, f);
  }

  FieldData *field(const std::string &fname) {
    auto it = fields.find(fname);
    return it == fields.end() ? nullptr : it->second.get();
  }
  const FieldData *field(const std::string &fname) const {
    auto it = fields.find(fname);
    return it == fields.end() ? nullptr : it->second.get();
  }
  const std::vector<std::string> &fieldNames() const { return order; }

  const int *getEndpointsData() const { return endpoints.data(); }
  int *getEndpointsData() { return endpoints.data(); }
  ElementRef getEndpoint(ElementRef e, int i) const {
    return ElementRef(endpoints[(size_t)e.getIdent() * getCardinality() + i]);
  }

  // element iteration: for (ElementRef e : set)
  class Iterator {
   public:
    explicit Iterator(int i) : i(i) {}
    ElementRef operator*() const { return ElementRef(i); }
    Iterator &operator++() { ++i; return *this; }
    bool operator!=(const Iterator &o) const { return i != o.i; }

   private:
    int i;
  };
  Iterator begin() const { return Iterator(0); }
  Iterator end() const { return Iterator(numElements); }

  // bumped whenever elements are added or reordered (a bound function must re-init)
  uint64_t structureVersion() const { return version; }

  // the 3-vector double field holding positions, used by reorder()
  // (graph.h:255-262 setSpatialField / getSpatialFieldName)
  void setSpatialField(const std::string &fname) {
    const FieldData *f = field(fname);
    if (!f || f->type != FieldType::Float64 || f->ncomp != 3)
      throw SimitException("spatial field " + fname + " must be a double 3-vector field");
    spatial = fname;
  }
  bool hasSpatialField() const { return !spatial.empty(); }
  const std::string &getSpatialFieldName() const { return spatial; }

  // Move element i to position newPos[i] (a permutation): every field's
  // values and, for an edge set, its endpoint rows.
  void permuteElements(const std::vector<int> &newPos) {
    if (kind == Grid) throw SimitException("a grid edge set's elements keep the grid's order");
    checkPermutation(newPos, numElements);
    for (auto &kv : fields) {
      FieldData &f = *kv.second;
      const size_t eb = f.elementBytes();
      std::vector<unsigned char> nb(f.bytes.size());
      for (int i = 0; i < numElements; ++i)
        std::memcpy(nb.data() + (size_t)newPos[i] * eb, f.bytes.data() + (size_t)i * eb, eb);
      f.bytes.swap(nb);
    }
    const int card = getCardinality();
    if (card) {
      std::vector<int> ne(endpoints.size());
      for (int i = 0; i < numElements; ++i)
        for (int c = 0; c < card; ++c)
          ne[(size_t)newPos[i] * card + c] = endpoints[(size_t)i * card + c];
      endpoints.swap(ne);
    }
    ++version;
  }
  // After the endpoint set `vertices` was permuted by newPos: renumber the
  // endpoints that refer to it.
  void remapEndpoints(const Set &vertices, const std::vector<int> &newPos) {
    checkPermutation(newPos, vertices.getSize());
    const int card = getCardinality();
    for (int c = 0; c < card; ++c) {
      if (endpointSets[c] != &vertices) continue;
      for (int i = 0; i < numElements; ++i) {
        int &e = endpoints[(size_t)i * card + c];
        e = newPos[e];
      }
    }
    ++version;
  }

 private:
  ElementRef grow() {
    for (auto &kv : fields) kv.second->bytes.resize(kv.second->bytes.size() + kv.second->elementBytes(), 0);
    ++version;
    return ElementRef(numElements++);
  }

  int64_t gridIndex(const std::vector<int> &coords, const char *what) const {
    if (kind != Grid)
      throw SimitException(std::string("Cannot retrieve ") + what + " of non-grid set");
    if (coords.size() != dimensions.size())
      throw SimitException("Must provide number of coords equal to the number of dimensions");
    int64_t index = 0, total = 1;
    for (int i = (int)dimensions.size() - 1; i >= 0; --i) {
      index = index * dimensions[i] + coords[i];
      total *= dimensions[i];
    }
    if (index < 0 || index >= total)
      throw SimitException("Coordinates must not be negative and must fall within the grid "
                           "dimensions");
    return index;
  }

  std::string name;
  Kind kind = Unstructured;
  std::vector<int> dimensions;
  int numElements = 0;
  uint64_t version = 0;
  std::vector<const Set *> endpointSets;
  std::vector<int> endpoints;
  std::map<std::string, std::unique_ptr<FieldData>> fields;
  std::vector<std::string> order;
  std::string spatial;

  static void checkPermutation(const std::vector<int> &p, int n) {
    if ((int)p.size() != n) throw SimitException("reordering: the map must have one entry per element");
    std::vector<char> seen(n, 0);
    for (int v : p) {
      if (v < 0 || v >= n || seen[v]) throw SimitException("reordering: the map is not a permutation");
      seen[v] = 1;
    }
  }
};

// createBox (src/graph.cpp:66-115): numX*numY*numZ points numbered
// x*numY*numZ + y*numZ + z, then x-edges, y-edges and z-edges.
class Box {
 public:
  Box(unsigned nx, unsigned ny, unsigned nz, std::vector<ElementRef> refs)
      : nX(nx), nY(ny), nZ(nz), refs(std::move(refs)) {}
  ElementRef operator()(unsigned x, unsigned y, unsigned z) const {
    return refs[(size_t)x * nY * nZ + (size_t)y * nZ + z];
  }
  unsigned numX() const { return nX; }
  unsigned numY() const { return nY; }
  unsigned numZ() const { return nZ; }

 private:
  unsigned nX, nY, nZ;
  std::vector<ElementRef> refs;
};
Box createBox(Set *vertices, Set *edges, unsigned numX, unsigned numY, unsigned numZ);

}  // namespace simit
