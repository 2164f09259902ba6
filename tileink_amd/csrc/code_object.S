/* Embeds the gfx950 code object (built from kernels.hip) into libtileink_hip.so;
 * tk_context_create hands tk_code_object_begin to hipModuleLoadData. */
    .section .rodata
    .balign 4096
    .globl tk_code_object_begin
tk_code_object_begin:
    .incbin TK_HSACO_PATH
    .globl tk_code_object_end
tk_code_object_end:
    .byte 0
    .section .note.GNU-stack,"",@progbits
