// Prints each double given in hexadecimal floating-point notation on stdin
// with std::ostream's default format, one per line: the number format of
// cpp_dmc/main.cpp's OBJ writer.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

int main() {
    std::string s;
    while (std::cin >> s) std::cout << std::strtod(s.c_str(), nullptr) << std::endl;
    return 0;
}
