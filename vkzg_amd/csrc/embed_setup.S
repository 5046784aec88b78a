/* The Ethereum mainnet G1 monomial trusted setup (4096 compressed points),
 * linked into the library so vkzg_prover_new needs no file
 * (reference src/setup_data.h:8-16 embeds the same bytes).
 * VKZG_SETUP_FILE is the quoted absolute path of vkzg_amd/data/g1_monomial_mainnet.bin. */
    .section .rodata
    .global vkzg_amd_mainnet_g1
    .balign 64
vkzg_amd_mainnet_g1:
    .incbin VKZG_SETUP_FILE
    .global vkzg_amd_mainnet_g1_end
vkzg_amd_mainnet_g1_end:
    .section .note.GNU-stack,"",@progbits
