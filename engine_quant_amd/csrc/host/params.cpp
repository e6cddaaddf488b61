// Params accessors, cpp/engine/src/params.cpp (same messages and exception types).
#include "engine.hpp"

namespace engine {

namespace {
template <class T>
const T* find_typed(const Params& params, const std::string& key, const char* type_name, bool required) {
    auto it = params.find(key);
    if (it == params.end()) {
        if (required) throw std::out_of_range("Params: falta la clave requerida '" + key + "'");
        return nullptr;
    }
    if (const T* v = std::get_if<T>(&it->second)) return v;
    throw std::invalid_argument("Params: la clave '" + key + "' no es un " + type_name);
}
}  // namespace

double get_double(const Params& params, const std::string& key) {
    return *find_typed<double>(params, key, "double", true);
}

double get_double(const Params& params, const std::string& key, double default_value) {
    const double* v = find_typed<double>(params, key, "double", false);
    return v ? *v : default_value;
}

bool get_bool(const Params& params, const std::string& key, bool default_value) {
    const bool* v = find_typed<bool>(params, key, "bool", false);
    return v ? *v : default_value;
}

const std::vector<double>& get_vector(const Params& params, const std::string& key) {
    return *find_typed<std::vector<double>>(params, key, "vector<double>", true);
}

const std::string& get_string(const Params& params, const std::string& key) {
    return *find_typed<std::string>(params, key, "string", true);
}

std::string get_string(const Params& params, const std::string& key, std::string default_value) {
    const std::string* v = find_typed<std::string>(params, key, "string", false);
    return v ? *v : default_value;
}

bool contains(const Params& params, const std::string& key) { return params.count(key) > 0; }

}  // namespace engine
