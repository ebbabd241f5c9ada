#!/bin/sh
# Writes tests/golden/chacha20_openssl.txt: ChaCha20 keystreams produced by
# OpenSSL (`openssl enc -chacha20`, IV = 32-bit little-endian block counter
# followed by a 96-bit nonce) for a few (key, counter) pairs. They pin the
# oracle's ChaCha block function independently of its own code; ChaCha8 is the
# same routine with 4 instead of 10 double rounds.
set -e
cd "$(dirname "$0")"
out=chacha20_openssl.txt
: > "$out"
tmp=$(mktemp)
head -c 256 /dev/zero > "$tmp"
for spec in \
  "0000000000000000000000000000000000000000000000000000000000000000 00000000" \
  "0000000000000000000000000000000000000000000000000000000000000001 01000000" \
  "000102030405060708090a0b0c0d0e0f101112131415161718191a1b1c1d1e1f 00000000" \
  "c0ffee00deadbeef0123456789abcdeffedcba98765432100f1e2d3c4b5a6978 05000000" ; do
  key=${spec% *}; ctr=${spec#* }
  ks=$(openssl enc -chacha20 -K "$key" -iv "${ctr}000000000000000000000000" -in "$tmp" | od -An -tx1 -v | tr -d ' \n')
  echo "$key $ctr $ks" >> "$out"
done
rm -f "$tmp"
