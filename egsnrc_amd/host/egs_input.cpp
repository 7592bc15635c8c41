// egs_input.cpp -- .egsinp tree parser (see egs_input.h).
#include "egs_input.h"

#include <cctype>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace egshost {

static std::string norm(const std::string &s) {
    std::string r;
    for (char c : s)
        if (!isspace((unsigned char)c) && c != '_') r += (char)tolower((unsigned char)c);
    return r;
}

bool InputItem::same(const std::string &a, const std::string &b) { return norm(a) == norm(b); }

const InputItem *InputItem::block(const std::string &name) const {
    for (auto &c : children)
        if (c->is_block && same(c->key, name)) return c.get();
    return nullptr;
}

std::vector<const InputItem *> InputItem::blocks(const std::string &name) const {
    std::vector<const InputItem *> r;
    for (auto &c : children)
        if (c->is_block && same(c->key, name)) r.push_back(c.get());
    return r;
}

const InputItem *InputItem::item(const std::string &name) const {
    for (auto &c : children)
        if (!c->is_block && same(c->key, name)) return c.get();
    return nullptr;
}

static std::vector<std::string> split_values(const std::string &v) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : v) {
        if (isspace((unsigned char)c) || c == ',' || c == ';') {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

int InputItem::get(const std::string &name, std::string &v) const {
    const InputItem *i = item(name);
    if (!i) return 1;
    v = i->value;
    return 0;
}

int InputItem::get(const std::string &name, double &v) const {
    std::vector<double> a;
    if (get(name, a) || a.size() != 1) return 1;
    v = a[0];
    return 0;
}

int InputItem::get(const std::string &name, int &v) const {
    double d;
    if (get(name, d)) return 1;
    v = (int)d;
    return 0;
}

int InputItem::get(const std::string &name, std::vector<double> &v) const {
    const InputItem *i = item(name);
    if (!i) return 1;
    v.clear();
    try {
        for (auto &t : split_values(i->value)) v.push_back(std::stod(t));
    } catch (...) {
        return 2;
    }
    return v.empty() ? 1 : 0;
}

int InputItem::get(const std::string &name, std::vector<int> &v) const {
    std::vector<double> d;
    int e = get(name, d);
    if (e) return e;
    v.clear();
    for (double x : d) v.push_back((int)x);
    return 0;
}

int InputItem::get(const std::string &name, std::vector<std::string> &v) const {
    const InputItem *i = item(name);
    if (!i) return 1;
    v = split_values(i->value);
    return v.empty() ? 1 : 0;
}

static std::string strip_comments(const std::string &text) {
    std::string out;
    size_t i = 0, n = text.size();
    while (i < n) {
        if (text.compare(i, 2, "/*") == 0) {
            size_t e = text.find("*/", i + 2);
            i = e == std::string::npos ? n : e + 2;
        } else if (text[i] == '#' || text.compare(i, 2, "//") == 0) {
            while (i < n && text[i] != '\n') i++;
        } else {
            out += text[i++];
        }
    }
    return out;
}

static std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

std::unique_ptr<InputItem> parse_input(const std::string &text) {
    auto root = std::make_unique<InputItem>();
    root->is_block = true;
    std::vector<InputItem *> st{root.get()};
    std::istringstream in(strip_comments(text));
    std::string line;
    while (std::getline(in, line)) {
        std::string t = trim(line);
        if (t.empty()) continue;
        if (t[0] == ':') {
            size_t e = t.find(':', 1);
            if (e == std::string::npos) throw std::runtime_error("malformed block line: " + t);
            std::string tag = trim(t.substr(1, e - 1));
            std::string low = norm(tag);
            if (low.compare(0, 5, "start") == 0) {
                auto b = std::make_unique<InputItem>();
                b->is_block = true;
                b->key = trim(tag.substr(5));
                InputItem *raw = b.get();
                st.back()->children.push_back(std::move(b));
                st.push_back(raw);
            } else if (low.compare(0, 4, "stop") == 0) {
                std::string name = trim(tag.substr(4));
                if (st.size() < 2 || !InputItem::same(st.back()->key, name))
                    throw std::runtime_error("unbalanced :stop " + name + ":");
                st.pop_back();
            } else {
                throw std::runtime_error("unknown block tag: " + t);
            }
            continue;
        }
        size_t eq = t.find('=');
        if (eq == std::string::npos) continue;
        auto it = std::make_unique<InputItem>();
        it->key = trim(t.substr(0, eq));
        it->value = trim(t.substr(eq + 1));
        st.back()->children.push_back(std::move(it));
    }
    if (st.size() != 1) throw std::runtime_error("missing :stop: for block " + st.back()->key);
    return root;
}

std::unique_ptr<InputItem> parse_input_file(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open input file " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_input(ss.str());
}

}  // namespace egshost
