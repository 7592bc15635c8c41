// egs_input.h -- the .egsinp input tree used by the host application.
//
// Mirrors the parts of EGS_Input (egs++/egs_input.h:216-334) the voxel dose
// path needs: nested ":start name:" / ":stop name:" blocks, "key = value"
// items, case-insensitive key comparison with blanks ignored (as
// EGS_Input::compare), "#" and "//" line comments and C-style comments.
#pragma once
#include <memory>
#include <string>
#include <vector>

namespace egshost {

struct InputItem {
    std::string key;                 // item key or block name
    std::string value;               // "key = value" text (blocks: empty)
    std::vector<std::unique_ptr<InputItem>> children;
    bool is_block = false;

    // key comparison of EGS_Input::compare: case-insensitive, blanks
    // and underscores ignored
    static bool same(const std::string &a, const std::string &b);

    const InputItem *block(const std::string &name) const;
    std::vector<const InputItem *> blocks(const std::string &name) const;
    const InputItem *item(const std::string &name) const;

    // getInput analogues: 0 = found, non-zero = missing/malformed
    int get(const std::string &name, std::string &v) const;
    int get(const std::string &name, double &v) const;
    int get(const std::string &name, int &v) const;
    int get(const std::string &name, std::vector<double> &v) const;
    int get(const std::string &name, std::vector<int> &v) const;
    int get(const std::string &name, std::vector<std::string> &v) const;
};

// Parse a whole input file; throws std::runtime_error on unbalanced blocks.
std::unique_ptr<InputItem> parse_input(const std::string &text);
std::unique_ptr<InputItem> parse_input_file(const std::string &path);

}  // namespace egshost
