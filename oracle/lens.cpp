int ob_lens_placeholder(void){return 0;}
