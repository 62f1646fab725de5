#ifndef _OUTPUT_DIR_H
#define _OUTPUT_DIR_H

#include <string>

std::string output_dir = "/root/repo/oracle/_ref/avx512_f16" ;

#endif
