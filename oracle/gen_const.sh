#!/bin/sh
# Writes FAME's CONST.h with the compile-time constants a configuration needs
# (the documented way to configure FAME: edit CONST.h and rebuild).
# usage: gen_const.sh <CONST.h> HTAB READLEN CORENUM CHROMNUM MINPD MAXPD CHUNK [KMERCUTOFF]
set -e
src=$1
cut=${9:-1500}
sed -e "s/^constexpr unsigned int READLEN = [0-9]*;/constexpr unsigned int READLEN = $3;/" \
    -e "s/^#define CORENUM [0-9]*/#define CORENUM $4/" \
    -e "s/^constexpr uint32_t MINPDIST = [0-9]*;/constexpr uint32_t MINPDIST = $6;/" \
    -e "s/^constexpr uint32_t MAXPDIST = [0-9]*;/constexpr uint32_t MAXPDIST = $7;/" \
    -e "s/^constexpr unsigned int CHROMNUM = [0-9]*;/constexpr unsigned int CHROMNUM = $5;/" \
    -e "s/^constexpr unsigned int CHUNKSIZE = [0-9]*;/constexpr unsigned int CHUNKSIZE = $8;/" \
    -e "s/^constexpr uint64_t HTABSIZE = .*;/constexpr uint64_t HTABSIZE = ${2}ULL;/" \
    -e "s/^constexpr uint64_t KMERCUTOFF = [0-9]*;/constexpr uint64_t KMERCUTOFF = $cut;/" \
    "$src"
